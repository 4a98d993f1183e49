// The merge of factor files into a candidate list (tools/ctools.cpp:151-248:
// read_cand, read_factors, write_cand): the host parses the files with the
// reference's skipping rules, the device verifies every pair whose b lies in
// the window (a grid-stride kernel over batches of pairs, 64-bit Montgomery
// arithmetic), and the host removes the b of every pair that holds from the
// sorted candidate vector and writes what remains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_numth.h"

namespace ct {

// ok[i] = 1 when p[i] divides b^(2^m) - b^(2^(m-1)) + 1 (mode +1) or - 1
// (mode -1): x = b^(2^(m-1)), x^2 - x = -1 or +1 (ctools.cpp:200-202).  The
// Montgomery product r = hi - mulhi(lo q, p) (+ p) holds for every odd
// p < 2^64; an even p or p < 3 is not a factor.
__global__ void k_merge_verify(const uint64_t *p, const uint64_t *b, uint64_t count, int m, int mode, uint8_t *ok)
{
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
	{
		const uint64_t pi = p[i];
		if ((pi & 1) == 0 || pi < 3) { ok[i] = 0; continue; }
		const m64d A(pi);
		uint64_t x = A.to_m(b[i]);
		for (int k = 0; k < m - 1; ++k) x = A.mul(x, x);
		const uint64_t r = A.sub(A.mul(x, x), x);
		ok[i] = (r == ((mode > 0) ? A.neg_one() : A.one)) ? 1 : 0;
	}
}

namespace {

// Unsigned decimal integers of a text file, read front to back through a
// fixed buffer (factor files can be larger than memory): next() skips any
// non-digit bytes and returns the following number; a number above 2^64 - 1
// sets overflow.
class number_reader
{
public:
	explicit number_reader(const char *path) : f_(std::fopen(path, "rb")), buf_(1u << 24) {}
	~number_reader() { if (f_) std::fclose(f_); }
	number_reader(const number_reader &) = delete;
	number_reader &operator=(const number_reader &) = delete;
	bool ok() const { return f_ != nullptr; }
	bool error() const { return err_; }
	bool next(uint64_t &v, bool &overflow)
	{
		int c;
		while ((c = get()) >= 0 && (c < '0' || c > '9')) {}
		if (c < 0) return false;
		v = 0;
		overflow = false;
		for (; c >= '0' && c <= '9'; c = get())
		{
			const uint64_t d = (uint64_t)(c - '0');
			if (v > (~0ull - d) / 10) overflow = true;
			v = v * 10 + d;
		}
		return true;
	}

private:
	int get()
	{
		if (pos_ == len_)
		{
			if (!f_ || eof_) return -1;
			len_ = std::fread(buf_.data(), 1, buf_.size(), f_);
			pos_ = 0;
			if (len_ < buf_.size()) { eof_ = true; err_ = std::ferror(f_) != 0; }
			if (len_ == 0) return -1;
		}
		return (unsigned char)buf_[pos_++];
	}
	FILE *f_;
	std::vector<char> buf_;
	size_t pos_ = 0, len_ = 0;
	bool eof_ = false, err_ = false;
};

constexpr uint64_t kBatch = 1ull << 22;       // pairs per device batch
constexpr uint64_t kTera = 1000000000000ull;

}   // namespace
}   // namespace ct

using namespace ct;

int ct_merge_factors(int device, int m, uint64_t b_min, uint64_t b_max, const char *cand_path, int nfiles,
                     const char *const *factor_paths, const int *modes, const uint64_t *p_min_t,
                     const uint64_t *p_max_t, const char *out_path, ct_merge_stats *stats)
{
	if (!cand_path || !out_path || nfiles < 0 || (nfiles > 0 && (!factor_paths || !modes)))
		return fail(CT_ERR_ARG, "null argument");
	if (m < 1 || m > 31) return fail(CT_ERR_ARG, "m must be in [1, 31]");
	for (int f = 0; f < nfiles; ++f)
	{
		if (!factor_paths[f]) return fail(CT_ERR_ARG, "null factor path");
		if (modes[f] != 1 && modes[f] != -1) return fail(CT_ERR_ARG, "mode must be +1 or -1");
	}
	ct_merge_stats st{};
	// read_cand (ctools.cpp:151-175): a b not above the previous one is skipped
	std::vector<uint64_t> cand;
	{
		number_reader in(cand_path);
		if (!in.ok()) return fail(CT_ERR_ARG, "cannot read the candidate file");
		uint64_t v = 0, prev = 0;
		bool ovf = false;
		while (in.next(v, ovf))
		{
			if (ovf) return fail(CT_ERR_ARG, "candidate above 2^64 - 1");
			if (v > prev) cand.push_back(v);
			prev = v;
		}
		if (in.error()) return fail(CT_ERR_ARG, "read error in the candidate file");
	}
	st.candidates = cand.size();
	std::vector<uint8_t> removed(cand.size(), 0);

	int rc = CT_OK;
	uint64_t *d_p = nullptr, *d_b = nullptr;
	uint8_t *d_ok = nullptr;
	hipStream_t hs = nullptr;
	std::vector<uint64_t> bp, bb;
	std::vector<uint8_t> ok;
	bp.reserve(1u << 16); bb.reserve(1u << 16);
	// device verification of the pending batch, then the removals
	auto flush = [&](int mode) -> int {
		if (bp.empty()) return CT_OK;
		const uint64_t cnt = bp.size();
		hipError_t h = hipSuccess;
		if (!hs)
		{   // the device is touched only when some pair lies in the window
			h = hipSetDevice(device);
			if (h == hipSuccess) h = hipStreamCreate(&hs);
			if (h == hipSuccess) h = hipMalloc((void **)&d_p, kBatch * sizeof(uint64_t));
			if (h == hipSuccess) h = hipMalloc((void **)&d_b, kBatch * sizeof(uint64_t));
			if (h == hipSuccess) h = hipMalloc((void **)&d_ok, kBatch);
			if (h != hipSuccess) return hip_fail(h, "merge: device buffers");
		}
		h = hipMemcpyAsync(d_p, bp.data(), cnt * sizeof(uint64_t), hipMemcpyHostToDevice, hs);
		if (h == hipSuccess) h = hipMemcpyAsync(d_b, bb.data(), cnt * sizeof(uint64_t), hipMemcpyHostToDevice, hs);
		if (h != hipSuccess) return hip_fail(h, "merge: copy of the pairs");
		const uint32_t grid = (uint32_t)std::min<uint64_t>((cnt + 255) / 256, 4096);
		hipLaunchKernelGGL(k_merge_verify, dim3(grid), dim3(256), 0, hs, d_p, d_b, cnt, m, mode, d_ok);
		h = hipGetLastError();
		if (h != hipSuccess) return hip_fail(h, "merge: verification launch");
		ok.resize(cnt);
		h = hipMemcpyAsync(ok.data(), d_ok, cnt, hipMemcpyDeviceToHost, hs);
		if (h == hipSuccess) h = hipStreamSynchronize(hs);
		if (h != hipSuccess) return hip_fail(h, "merge: verification");
		for (uint64_t i = 0; i < cnt; ++i)
		{
			if (!ok[i]) { ++st.bad; continue; }
			++st.factors;
			const auto it = std::lower_bound(cand.begin(), cand.end(), bb[i]);
			if (it != cand.end() && *it == bb[i]) removed[(size_t)(it - cand.begin())] = 1;
		}
		bp.clear(); bb.clear();
		return CT_OK;
	};
	// read_factors (ctools.cpp:177-224), one file after the other
	for (int f = 0; f < nfiles && rc == CT_OK; ++f)
	{
		number_reader in(factor_paths[f]);
		if (!in.ok()) { rc = fail(CT_ERR_ARG, "cannot read a factor file"); break; }
		const bool ranged = p_min_t && p_max_t;
		const unsigned __int128 lo = ranged ? (unsigned __int128)p_min_t[f] * kTera : 0;
		const unsigned __int128 hi = ranged ? (unsigned __int128)p_max_t[f] * kTera : 0;
		uint64_t p = 0, b = 0, p_prev = 0, b_prev = 0;
		bool ovf = false;
		while (true)
		{
			if (!in.next(p, ovf)) break;
			if (ovf) { rc = fail(CT_ERR_ARG, "factor above 2^64 - 1"); break; }
			if (!in.next(b, ovf)) { rc = fail(CT_ERR_ARG, "factor file line without b"); break; }
			if (ovf) { rc = fail(CT_ERR_ARG, "b above 2^64 - 1"); break; }
			if (p == p_prev && b == b_prev) continue;
			p_prev = p; b_prev = b;
			++st.pairs;
			if (ranged && ((unsigned __int128)p < lo || (unsigned __int128)p > hi)) ++st.out_of_range;
			if (b < b_min || b > b_max) continue;
			++st.in_window;
			bp.push_back(p); bb.push_back(b);
			if (bp.size() == kBatch && (rc = flush(modes[f])) != CT_OK) break;
		}
		if (rc == CT_OK && in.error()) rc = fail(CT_ERR_ARG, "read error in a factor file");
		if (rc == CT_OK) rc = flush(modes[f]);
	}
	// write_cand (ctools.cpp:226-248)
	if (rc == CT_OK)
	{
		FILE *out = std::fopen(out_path, "wb");
		if (!out) { rc = fail(CT_ERR_ARG, "cannot write the candidate file"); goto done; }
		std::string buf;
		buf.reserve(1 << 20);
		char tmp[24];
		bool wrote = true;
		for (size_t i = 0; i < cand.size(); ++i)
		{
			if (removed[i]) continue;
			++st.remaining;
			const int len = std::snprintf(tmp, sizeof(tmp), "%llu\n", (unsigned long long)cand[i]);
			buf.append(tmp, (size_t)len);
			if (buf.size() >= (1u << 20))
			{
				wrote = wrote && std::fwrite(buf.data(), 1, buf.size(), out) == buf.size();
				buf.clear();
			}
		}
		wrote = wrote && std::fwrite(buf.data(), 1, buf.size(), out) == buf.size();
		if (std::fclose(out) != 0 || !wrote) rc = fail(CT_ERR_ARG, "short write of the candidate file");
	}
done:
	if (d_p) (void)hipFree(d_p);
	if (d_b) (void)hipFree(d_b);
	if (d_ok) (void)hipFree(d_ok);
	if (hs) (void)hipStreamDestroy(hs);
	if (stats) *stats = st;
	return rc;
}
