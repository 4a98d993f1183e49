// The reference's on-disk formats, written and read by the host:
//   - the sieve state file sv_<m>_<b_min>_<b_max>.dat and the candidate list
//     cand_<m>_<b_min>_<b_max>.txt (Sieve::write / Sieve::read,
//     sieve/CPU/ctsieve.cpp:551-629);
//   - the PRP checkpoint cyclo.ctx (cyclo::saveContext / _readContext,
//     cyclo/src/cyclo.h:107-176, transform::saveContext / readContext,
//     cyclo/src/transform.h:445-468, CRC of cyclo/src/file.h:14-121).
// Both end with the CRC-32 (IEEE 802.3 polynomial, reflected, 0xedb88320) of
// every preceding byte, stored as ~crc ^ 0xa23777ac; both rotate the previous
// file to <name>.old before the new one takes its name.
#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"

using namespace ct;

namespace {

constexpr uint32_t kCrcTag = 0xa23777acu;
// the reference engine's primes (cyclo/src/transform.h:26-28): the checkpoint
// stores each carried digit as its residues modulo these
constexpr uint32_t kRefP1 = 4293918721u, kRefP2 = 4253024257u, kRefP3 = 4231004161u;

struct crc32_ieee
{
	uint32_t table[256];
	crc32_ieee()
	{
		for (uint32_t i = 0; i < 256; ++i)
		{
			uint32_t r = i;
			for (int j = 0; j < 8; ++j) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : (r >> 1);
			table[i] = r;
		}
	}
	// running CRC: crc(previous, more bytes)
	uint32_t update(uint32_t crc, const void *buf, size_t len) const
	{
		const uint8_t *p = static_cast<const uint8_t *>(buf);
		uint32_t c = ~crc;
		for (size_t i = 0; i < len; ++i) c = (c >> 8) ^ table[(c ^ p[i]) & 0xff];
		return ~c;
	}
};

const crc32_ieee &crc_table() { static const crc32_ieee t; return t; }

bool file_exists(const std::string &f) { struct stat s; return stat(f.c_str(), &s) == 0; }

// Writes `bytes` to name.new, then name -> name.old, name.new -> name.
int rotate_write(const std::string &name, const std::vector<uint8_t> &bytes)
{
	const std::string tmp = name + ".new", old = name + ".old";
	FILE *f = std::fopen(tmp.c_str(), "wb");
	if (!f) return fail(CT_ERR_ARG, "cannot create the file");
	const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
	if (std::fclose(f) != 0 || !ok) return fail(CT_ERR_ARG, "cannot write the file");
	std::remove(old.c_str());
	if (file_exists(name) && std::rename(name.c_str(), old.c_str()) != 0) return fail(CT_ERR_ARG, "cannot rotate the old file");
	if (std::rename(tmp.c_str(), name.c_str()) != 0) return fail(CT_ERR_ARG, "cannot rename the new file");
	return CT_OK;
}

bool read_all(const std::string &name, std::vector<uint8_t> &out)
{
	FILE *f = std::fopen(name.c_str(), "rb");
	if (!f) return false;
	out.clear();
	uint8_t buf[1 << 16];
	size_t r;
	while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + r);
	std::fclose(f);
	return true;
}

template <typename T> void put(std::vector<uint8_t> &v, const T &x)
{
	const uint8_t *p = reinterpret_cast<const uint8_t *>(&x);
	v.insert(v.end(), p, p + sizeof(T));
}

template <typename T> T get(const std::vector<uint8_t> &v, size_t off) { T x; std::memcpy(&x, &v[off], sizeof(T)); return x; }

void seal(std::vector<uint8_t> &v)
{
	const uint32_t c = ~crc_table().update(0, v.data(), v.size()) ^ kCrcTag;
	put(v, c);
}

bool sealed(const std::vector<uint8_t> &v)
{
	if (v.size() < 4) return false;
	const uint32_t want = ~crc_table().update(0, v.data(), v.size() - 4) ^ kCrcTag;
	return get<uint32_t>(v, v.size() - 4) == want;
}

// <dir>/<prefix>_<m>_<b_min>_<b_max><suffix> (the _filename of ctsieve.cpp:477-478)
std::string sieve_name(const char *dir, const char *prefix, int m, uint64_t b_min, uint64_t b_max, const char *suffix)
{
	const std::string leaf = std::string(prefix) + "_" + std::to_string(m) + "_" + std::to_string(b_min) + "_" +
	                         std::to_string(b_max) + suffix;
	return (dir && *dir) ? std::string(dir) + "/" + leaf : leaf;
}

uint32_t seti(int32_t d, uint32_t p) { return (d < 0) ? (uint32_t)((int64_t)d + p) : (uint32_t)d; }

}   // namespace

uint32_t ct_crc32(uint32_t crc, const void *buf, size_t len) { return crc_table().update(crc, buf, len); }

int ct_sieve_save(const char *dir, int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_min_neg,
                  const uint8_t *struck, int write_cand)
{
	if (!struck || b_max < b_min) return fail(CT_ERR_ARG, "bad argument");
	const std::string sv = sieve_name(dir, "sv", m, b_min, b_max, ".dat");
	const std::string cand = sieve_name(dir, "cand", m, b_min, b_max, ".txt");
	const uint64_t size = b_max - b_min + 1;
	std::vector<uint32_t> idx;
	for (uint64_t i = 0; i < size; ++i) if (!struck[i]) idx.push_back((uint32_t)i);
	std::vector<uint8_t> v;
	v.reserve(24 + 4 * idx.size() + 4);
	put(v, p_min_pos); put(v, p_min_neg); put(v, (uint64_t)idx.size());
	for (uint32_t i : idx) put(v, i);
	seal(v);
	int rc = rotate_write(sv, v);
	if (rc || !write_cand) return rc;
	FILE *f = std::fopen(cand.c_str(), "w");
	if (!f) return fail(CT_ERR_ARG, "cannot create the candidate file");
	for (uint32_t i : idx) std::fprintf(f, "%llu\n", (unsigned long long)(b_min + i));
	return (std::fclose(f) == 0) ? CT_OK : fail(CT_ERR_ARG, "cannot write the candidate file");
}

int ct_sieve_load(const char *dir, int m, uint64_t b_min, uint64_t b_max, uint64_t *p_min_pos, uint64_t *p_min_neg,
                  uint8_t *struck)
{
	if (!struck || !p_min_pos || !p_min_neg || b_max < b_min) return fail(CT_ERR_ARG, "bad argument");
	const std::string sv = sieve_name(dir, "sv", m, b_min, b_max, ".dat");
	std::vector<uint8_t> v;
	if (!read_all(sv, v) || v.size() < 28) return fail(CT_ERR_ARG, "cannot read the sieve file");
	const uint64_t count = get<uint64_t>(v, 16), size = b_max - b_min + 1;
	if (v.size() != 24 + 4 * count + 4 || !sealed(v)) return fail(CT_ERR_ARG, "invalid sieve file (size or crc32)");
	std::memset(struck, 1, size);
	for (uint64_t c = 0; c < count; ++c)
	{
		const uint32_t i = get<uint32_t>(v, 24 + 4 * c);
		if (i >= size) return fail(CT_ERR_ARG, "invalid sieve file (index beyond the window)");
		struck[i] = 0;
	}
	*p_min_pos = get<uint64_t>(v, 0);
	*p_min_neg = get<uint64_t>(v, 8);
	return CT_OK;
}

int ct_context_save(ct_engine *e, const char *path, int k, int i, double elapsed)
{
	if (!e || !path) return fail(CT_ERR_ARG, "bad argument");
	const size_t L = e->lanes, n = e->n, size = 2 * L * n;
	std::vector<int32_t> d0(L * n), d1(L * n);
	int rc = ct_engine_read_state(e, 0, d0.data());
	if (rc == CT_OK) rc = ct_engine_read_state(e, 1, d1.data());
	if (rc) return rc;
	std::vector<uint8_t> v;
	v.reserve(20 + size * 12 + 4);
	put(v, (int32_t)1); put(v, (int32_t)k); put(v, (int32_t)i); put(v, elapsed);
	// registers 0 (x) and 1 (d(t)): index (reg n + j) L + lane, as the
	// reference's device buffers x12 (two residues) then x3
	std::vector<uint32_t> x12(2 * size), x3(size);
	for (size_t r = 0; r < 2; ++r)
	{
		const std::vector<int32_t> &d = r ? d1 : d0;
		for (size_t l = 0; l < L; ++l)
			for (size_t j = 0; j < n; ++j)
			{
				const int32_t x = d[l * n + j];
				const size_t at = (r * n + j) * L + l;
				x12[2 * at] = seti(x, kRefP1); x12[2 * at + 1] = seti(x, kRefP2); x3[at] = seti(x, kRefP3);
			}
	}
	const uint8_t *p12 = reinterpret_cast<const uint8_t *>(x12.data()), *p3 = reinterpret_cast<const uint8_t *>(x3.data());
	v.insert(v.end(), p12, p12 + x12.size() * 4);
	v.insert(v.end(), p3, p3 + x3.size() * 4);
	seal(v);
	return rotate_write(path, v);
}

static int context_try(ct_engine *e, const std::string &path, int *k, int *i, double *elapsed)
{
	std::vector<uint8_t> v;
	if (!read_all(path, v)) return -1;
	const size_t L = e->lanes, n = e->n, size = 2 * L * n;
	if (v.size() != 20 + size * 12 + 4 || get<int32_t>(v, 0) != 1 || !sealed(v)) return -2;
	std::vector<int32_t> d0(L * n), d1(L * n);
	const size_t o12 = 20, o3 = 20 + size * 8;
	for (size_t r = 0; r < 2; ++r)
		for (size_t l = 0; l < L; ++l)
			for (size_t j = 0; j < n; ++j)
			{
				const size_t at = (r * n + j) * L + l;
				const uint32_t r1 = get<uint32_t>(v, o12 + 8 * at), r2 = get<uint32_t>(v, o12 + 8 * at + 4);
				const uint32_t r3 = get<uint32_t>(v, o3 + 4 * at);
				// signed read of the third residue (transform.h:47, :526)
				const int32_t x = (r3 > kRefP3 / 2) ? (int32_t)(r3 - kRefP3) : (int32_t)r3;
				if (seti(x, kRefP1) != r1 || seti(x, kRefP2) != r2) return -2;      // not a carried state
				(r ? d1 : d0)[l * n + j] = x;
			}
	if (ct_engine_write_state(e, 0, d0.data()) || ct_engine_write_state(e, 1, d1.data())) return -3;
	*k = get<int32_t>(v, 4); *i = get<int32_t>(v, 8); *elapsed = get<double>(v, 12);
	return 0;
}

int ct_context_load(ct_engine *e, const char *path, int *k, int *i, double *elapsed)
{
	if (!e || !path || !k || !i || !elapsed) return fail(CT_ERR_ARG, "bad argument");
	int r = context_try(e, path, k, i, elapsed);
	if (r == -3) return fail(CT_ERR_HIP, "cannot write the restored registers");
	if (r < 0)
	{
		const int r_old = context_try(e, std::string(path) + ".old", k, i, elapsed);
		if (r_old == -3) return fail(CT_ERR_HIP, "cannot write the restored registers");
		if (r_old == 0) return CT_OK;
		return (r == -1 && r_old == -1) ? fail(CT_ERR_ARG, "no context file") : fail(CT_ERR_ARG, "invalid context file");
	}
	return CT_OK;
}
