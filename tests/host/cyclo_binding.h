// The transform binding of the integration guide, section 1: cyclo's
// `transform` (cyclo/src/transform.h:234-655) reimplemented over the C ABI of
// include/ctwin.h.  The guide prints this class; this file is the copy that
// is compiled and run (tests/test_binding.py).
#pragma once

#include "ctwin.h"

class transform {
    ct_engine *_e = nullptr;
    size_t _n; vint32 _b;
    std::vector<uint32> _x;                      // 3 host slots, as transform.h:241
    static void check(int rc) { if (rc) pio::error(ct_last_error(), true); }
public:
    transform(size_t size, engine &, bool, size_t csize)
        : _n(size), _x(3 * VSIZE * size)
    {   // csize 0 runs the autotuner (engine kind, column tile), as the reference tunes CSIZE
        check(ct_engine_create(/*device*/0, ilog2_32(uint32_t(size)), VSIZE, csize ? 16 : 0, CT_KIND_AUTO, &_e));
    }
    ~transform() { ct_engine_destroy(_e); }

    void set_b(const vint32 &b) { _b = b; check(ct_engine_set_b(_e, b.data())); }       // transform.h:472
    void set(uint32 a) const { check(ct_engine_set(_e, a)); }                          // engine.h:231
    void copy(uint32 dst, uint32 src) const { check(ct_engine_copy(_e, dst, src)); }   // engine.h:240
    void squareDup(uint64 dup) const { check(ct_engine_square_dup(_e, &dup)); }        // engine.h:387
    void mul() const { check(ct_engine_mul(_e)); }                                     // engine.h:421
    // getInt(reg) reads *device* register reg here; the reference always reads
    // device register 0 into host slot reg (transform.h:516), so call
    // getInt(0) where the reference reads reg 0, and mirror its copy(0, 2)
    // before getInt(2), exactly as cyclo.h:373-376 does.
    void getInt(uint32 reg) { check(ct_engine_get_int(_e, 0, &_x[reg * VSIZE * _n])); }
    bool isPrime(bool *prm, uint64_t *res64) const {                                   // transform.h:600
        int p[VSIZE]; const int rc = ct_engine_is_prime(_e, _x.data(), p, res64);
        for (size_t i = 0; i < VSIZE; ++i) prm[i] = p[i] != 0;
        return rc == CT_ERR_COMPUTE;
    }
    size_t get_csize() const { return 16; }
};
