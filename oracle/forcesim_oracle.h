/* Test infrastructure only: the CPU restatement of layout.forceatlas2.forcesim (SciGraphs'
 * ForceSim as the reference ports it, crates/graph-core/src/layout/force/forcesim*), the
 * checker the device path is compared against. See forcesim_oracle.c for the pins. */
#ifndef FORCESIM_ORACLE_H
#define FORCESIM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* numpy's default_rng(seed): the seeded PCG64 (state high, low, inc high, low), raw words
 * and 53-bit doubles */
void gro_pcg64_state(uint64_t seed, uint64_t out4[4]);
void gro_pcg64_raw(uint64_t seed, uint32_t count, uint64_t *out);
void gro_pcg64_doubles(uint64_t seed, uint32_t count, double *out);

/* numpy's reduction orders */
float gro_pairwise_f32(const float *v, size_t n);
double gro_pairwise_f64(const double *v, size_t n);
float gro_dot3_f32(const float *a, const float *b);

/* the correctly rounded cube root of an integer-valued x in [1, 2^53] */
double gro_cbrt_cr(double x);

/* k, repulsion, gravity (state.rs:117-132) for masses mass[0..n) */
void gro_forcesim_constants(uint32_t n, const float *mass, double scaling_ratio, double gravity,
                            double out3[3]);

/* The whole stage: simple graph, seeded start, iterations, rescale. out_xyz: n*3 f32
 * row-major; consts3 (may be NULL): k, repulsion, gravity. Returns 0, or 2 when a coordinate
 * is not finite (NonFinite{node.x}; the output is zeroed). */
int gro_forcesim_run(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m_raw,
                     uint32_t iterations, double scale, double scaling_ratio, double gravity,
                     double jitter_tolerance, uint64_t seed, float *out_xyz, double *consts3);

#ifdef __cplusplus
}
#endif
#endif
