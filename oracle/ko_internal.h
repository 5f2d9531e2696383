/* ko_internal.h -- shared types of the oracle (test infrastructure only). */
#ifndef KO_INTERNAL_H
#define KO_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

/* lanevector.rs:3-6 LaneVector<Avx512Lane, 9>: blocks of [9][8] + remainder. */
typedef struct {
    size_t nblocks;
    size_t nrem;
    double* blocks; /* nblocks * 72 */
    double rem[7 * 9];
} ko_lanevec;

#define KO_DECL_VARIANT(sfx)                                                                                    \
    double ko_lane_log_##sfx(double d);                                                                         \
    double ko_jacq_obj_lv_##sfx(const ko_lanevec* lv, const double x[9], double eps);                           \
    void ko_jacq_gh_lv_##sfx(const ko_lanevec* lv, const double x[9], double eps, double g[9], double h[81]);   \
    double ko_allele_obj2_##sfx(const double* blocks, size_t nb, const double* rem, size_t nrem,              \
                                const double x[2], double eps);                                                 \
    void ko_allele_gh2_##sfx(const double* blocks, size_t nb, const double* rem, size_t nrem, const double x[2], \
                             double eps, double g[2], double h[4]);                                          \
    void ko_gram_lv_##sfx(const ko_lanevec* lv, double n, double out[81]);

KO_DECL_VARIANT(avx512)
KO_DECL_VARIANT(avx2)
KO_DECL_VARIANT(base)

/* ko_lanes_fast.c: the REF loops in the reference's AVX-512 loop shape
 * (bit-identical to the variants above; AVX-512 hosts only). */
double ko_lane_log8_fast(double d, int lane);
double ko_jacq_obj_lv_fast(const ko_lanevec* lv, const double x[9], double eps);
void ko_jacq_gh_lv_fast(const ko_lanevec* lv, const double x[9], double eps, double g[9], double h[81]);
size_t ko_pair_lanes_codes_fast(const uint8_t* gx, const uint8_t* gy, const double* p, const double* q, size_t V,
                                ko_lanevec* lv);
int ko_has_avx512(void);

typedef struct {
    double (*log)(double);
    double (*jacq_obj)(const ko_lanevec*, const double*, double);
    void (*jacq_gh)(const ko_lanevec*, const double*, double, double*, double*);
    double (*allele_obj)(const double*, size_t, const double*, size_t, const double*, double);
    void (*allele_gh)(const double*, size_t, const double*, size_t, const double*, double, double*, double*);
    void (*gram)(const ko_lanevec*, double, double*);
    const char* name;
} ko_lane_ops;

const ko_lane_ops* ko_ops(void);

#endif
