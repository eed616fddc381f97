// C entry points over the reference's own distance routines, compiled from
// the reference sources in place (see Makefile, target _ref/libref_dist.so).
// TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle's restatement of
// VMath::l2norm (src/vmath.cpp:56-88) and DistOpt::L2SqrSIMDExt
// (src/distopt.cpp:11-80).
#include "vmath.h"
#include "distopt.h"

extern "C" float ref_l2norm(const float* v1, unsigned long s1, const float* v2,
                            unsigned long s2, unsigned d) {
    return VMath::l2norm(v1, s1, v2, s2, d);
}

extern "C" float ref_l2sqr_simd_ext(const float* v1, unsigned loc1, const float* v2,
                                    unsigned loc2, unsigned d) {
    alignas(16) float tmp[4];
    return DistOpt::L2SqrSIMDExt(v1, loc1, v2, loc2, d, tmp);
}

// DistOpt::l1norm (src/distopt.cpp:82-110), DistOpt::KSquare (:163-184) and
// DistOpt::cosine (:113-160): pin the oracle's restatements of the other
// metrics.
extern "C" float ref_l1norm(const float* v1, unsigned long s1, const float* v2, unsigned long s2, unsigned d) {
    return DistOpt::l1norm(v1, s1, v2, s2, d);
}

extern "C" float ref_ksquare(const float* v1, unsigned long s1, const float* v2, unsigned long s2, unsigned d) {
    return DistOpt::KSquare(v1, s1, v2, s2, d);
}

extern "C" float ref_cosine(const float* v1, unsigned s1, const float* v2, unsigned s2, unsigned d) {
    return DistOpt::cosine(v1, s1, v2, s2, d);
}
