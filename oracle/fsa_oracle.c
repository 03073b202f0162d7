/*
 * fsa_oracle.c -- CPU oracle for the full-state gate path (TEST INFRASTRUCTURE ONLY).
 *
 * Single-rank restatement of the reference's loops in C99 + OpenMP.  Every
 * function cites the reference lines it follows.  The oracle contract for the
 * reference's defects D1-D3 is stated in fsa_oracle.h.  Loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 */
#include "fsa_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cx;
typedef uint64_t idx_t;

/* ------------------------------------------------------------------------ */
/* index arithmetic (reference src/bit_maths.hpp:19-125)                      */
/* ------------------------------------------------------------------------ */

static inline idx_t bit_of(idx_t x, unsigned b) { return (x >> b) & 1ULL; }

/* open a gap at position b and put v there (bit_maths.hpp:43-49) */
static inline idx_t ins_bit(idx_t x, unsigned b, unsigned v) {
    idx_t low = x & ((1ULL << b) - 1ULL);
    return ((x >> b) << (b + 1)) | ((idx_t)v << b) | low;
}

/* insert v at every position of an ascending list (bit_maths.hpp:52-59) */
static inline idx_t ins_bits(idx_t x, const unsigned* asc, unsigned num, unsigned v) {
    for (unsigned i = 0; i < num; i++)
        x = ins_bit(x, asc[i], v);
    return x;
}

/* bit i of `value` -> position pos[i]; D1 corrected form of bit_maths.hpp:62-78 */
static inline idx_t put_bits(idx_t x, const unsigned* pos, unsigned num, idx_t value) {
    for (unsigned i = 0; i < num; i++) {
        idx_t m = 1ULL << pos[i];
        x = (x & ~m) | (bit_of(value, i) << pos[i]);
    }
    return x;
}

static inline unsigned parity64(idx_t x) { return (unsigned)__builtin_parityll(x); }

static inline idx_t mask_of(const unsigned* q, unsigned num) {
    idx_t m = 0;
    for (unsigned i = 0; i < num; i++) m ^= 1ULL << q[i];
    return m;
}

static void sort_asc(unsigned* q, unsigned num) {
    for (unsigned i = 1; i < num; i++) {
        unsigned v = q[i];
        int j = (int)i - 1;
        while (j >= 0 && q[j] > v) { q[j + 1] = q[j]; j--; }
        q[j + 1] = v;
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* initial states                                                            */
/* ------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform in (0,1) from counter c of stream `key` */
static inline double unit_open(uint64_t key, uint64_t c) {
    return ((double)(mix64(key + c) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* fixed-order blocked sum so the result does not depend on the thread count */
double orc_norm2(const orc_real* a, unsigned nq) {
    idx_t num = 1ULL << nq;
    idx_t blk = 4096, numBlk = (num + blk - 1) / blk;
    double* part = (double*)malloc(numBlk * sizeof(double));
    #pragma omp parallel for
    for (idx_t b = 0; b < numBlk; b++) {
        double s = 0;
        idx_t end = (b + 1) * blk < num ? (b + 1) * blk : num;
        for (idx_t j = b * blk; j < end; j++)
            s += a[2 * j] * a[2 * j] + a[2 * j + 1] * a[2 * j + 1];
        part[b] = s;
    }
    double s = 0;
    for (idx_t b = 0; b < numBlk; b++) s += part[b];
    free(part);
    return s;
}

/* i.i.d. N(0,1) real and imaginary parts by Box-Muller over a counter-based
 * splitmix64 stream, then normalised to unit 2-norm. */
void orc_init_gaussian(orc_real* a, unsigned nq, uint64_t seed) {
    idx_t num = 1ULL << nq;
    uint64_t key = mix64(seed);
    #pragma omp parallel for
    for (idx_t j = 0; j < num; j++) {
        double u1 = unit_open(key, 2 * j);
        double u2 = unit_open(key, 2 * j + 1);
        double r = sqrt(-2.0 * log(u1));
        double ph = 6.283185307179586476925286766559 * u2;
        a[2 * j] = r * cos(ph);
        a[2 * j + 1] = r * sin(ph);
    }
    double inv = 1.0 / sqrt(orc_norm2(a, nq));
    #pragma omp parallel for
    for (idx_t j = 0; j < 2 * num; j++) a[j] *= inv;
}

/* ------------------------------------------------------------------------ */
/* statevector (reference src/local_statevector.hpp)                          */
/* ------------------------------------------------------------------------ */

/* local_statevector.hpp:14-29 */
void orc_sv_oneTargGate(orc_real* raw, unsigned n, unsigned t, const orc_real* graw) {
    cx* amps = (cx*)raw;
    const cx* g = (const cx*)graw;
    cx g00 = g[0], g01 = g[1], g10 = g[2], g11 = g[3];
    idx_t numIts = (1ULL << n) / 2;
    #pragma omp parallel for
    for (idx_t j = 0; j < numIts; j++) {
        idx_t i0 = ins_bit(j, t, 0);
        idx_t i1 = i0 ^ (1ULL << t);
        cx a0 = amps[i0], a1 = amps[i1];
        amps[i0] = g00 * a0 + g01 * a1;
        amps[i1] = g10 * a0 + g11 * a1;
    }
}

/* local_statevector.hpp:32-51 */
void orc_sv_manyCtrlOneTargGate(orc_real* raw, unsigned n, const unsigned* ctrls, unsigned nc,
                                unsigned t, const orc_real* graw) {
    cx* amps = (cx*)raw;
    const cx* g = (const cx*)graw;
    cx g00 = g[0], g01 = g[1], g10 = g[2], g11 = g[3];
    unsigned q[64];
    for (unsigned i = 0; i < nc; i++) q[i] = ctrls[i];
    q[nc] = t;
    sort_asc(q, nc + 1);
    idx_t numIts = (1ULL << n) >> (nc + 1);
    #pragma omp parallel for
    for (idx_t j = 0; j < numIts; j++) {
        idx_t i1 = ins_bits(j, q, nc + 1, 1);
        idx_t i0 = i1 ^ (1ULL << t);
        cx a0 = amps[i0], a1 = amps[i1];
        amps[i0] = g00 * a0 + g01 * a1;
        amps[i1] = g10 * a0 + g11 * a1;
    }
}

/* local_statevector.hpp:54-69 */
void orc_sv_swapGate(orc_real* raw, unsigned n, unsigned qb1, unsigned qb2) {
    cx* amps = (cx*)raw;
    if (qb1 > qb2) { unsigned s = qb1; qb1 = qb2; qb2 = s; }
    if (qb1 == qb2) return;
    idx_t numIts = (1ULL << n) / 4;
    #pragma omp parallel for
    for (idx_t k = 0; k < numIts; k++) {
        idx_t j11 = ins_bit(ins_bit(k, qb1, 1), qb2, 1);
        idx_t j10 = j11 ^ (1ULL << qb1);
        idx_t j01 = j11 ^ (1ULL << qb2);
        cx tmp = amps[j01];
        amps[j01] = amps[j10];
        amps[j10] = tmp;
    }
}

/* local_statevector.hpp:72-99, with the D1-corrected setBits */
void orc_sv_manyTargGate(orc_real* raw, unsigned n, const unsigned* targets, unsigned k,
                         const orc_real* graw) {
    cx* amps = (cx*)raw;
    const cx* g = (const cx*)graw;
    unsigned sorted[64];
    for (unsigned i = 0; i < k; i++) sorted[i] = targets[i];
    sort_asc(sorted, k);
    idx_t dim = 1ULL << k;
    idx_t numOuterIts = (1ULL << n) / dim;
    idx_t* off = (idx_t*)malloc(dim * sizeof(idx_t));
    for (idx_t j = 0; j < dim; j++) off[j] = put_bits(0, targets, k, j);
    #pragma omp parallel
    {
        cx* cache = (cx*)malloc(dim * sizeof(cx));
        #pragma omp for
        for (idx_t o = 0; o < numOuterIts; o++) {
            idx_t base = ins_bits(o, sorted, k, 0);
            for (idx_t j = 0; j < dim; j++) cache[j] = amps[base | off[j]];
            for (idx_t j = 0; j < dim; j++) {
                cx acc = 0;
                for (idx_t l = 0; l < dim; l++) acc += g[j * dim + l] * cache[l];
                amps[base | off[j]] = acc;
            }
        }
        free(cache);
    }
    free(off);
}

/* distributed_statevector.hpp:244-276 (mask and powI preparation, 1-rank branch)
 * and local_statevector.hpp:102-135 (pair loop), with D1 and D2 corrected. */
static void pauli_tensor_or_gadget(cx* amps, unsigned n, const unsigned* targets,
                                   const unsigned* paulis, unsigned num, cx thisFac, cx otherFac) {
    cx powI = 1;
    unsigned xy[64];
    unsigned numXY = 0;
    idx_t maskXY = 0, maskYZ = 0;
    for (unsigned i = 0; i < num; i++) {
        if (paulis[i] == 2) powI *= I;
        if (paulis[i] == 1 || paulis[i] == 2) { xy[numXY++] = targets[i]; maskXY ^= 1ULL << targets[i]; }
        if (paulis[i] == 2 || paulis[i] == 3) maskYZ ^= 1ULL << targets[i];
    }
    idx_t numAmps = 1ULL << n;
    if (numXY == 0) {
        /* D2: the pair (j0, j0 ^ maskXY) degenerates to j0 itself */
        #pragma omp parallel for
        for (idx_t j = 0; j < numAmps; j++) {
            cx b = (1.0 - 2.0 * parity64(j & maskYZ)) * powI;
            cx a = amps[j];
            amps[j] = thisFac * a + otherFac * b * a;
        }
        return;
    }
    sort_asc(xy, numXY);
    idx_t numOuterIts = numAmps >> numXY;
    idx_t numInnerIts = (1ULL << numXY) / 2;
    #pragma omp parallel for
    for (idx_t k = 0; k < numOuterIts; k++) {
        idx_t h = ins_bits(k, xy, numXY, 0);
        for (idx_t l = 0; l < numInnerIts; l++) {
            idx_t j0 = put_bits(h, xy, numXY, l);
            idx_t j1 = j0 ^ maskXY;
            cx b0 = (1.0 - 2.0 * parity64(j0 & maskYZ)) * powI;
            cx b1 = (1.0 - 2.0 * parity64(j1 & maskYZ)) * powI;
            cx a0 = amps[j0], a1 = amps[j1];
            amps[j0] = thisFac * a0 + otherFac * b1 * a1;
            amps[j1] = thisFac * a1 + otherFac * b0 * a0;
        }
    }
}

/* distributed_statevector.hpp:279-284 */
void orc_sv_pauliTensor(orc_real* raw, unsigned n, const unsigned* targets, const unsigned* paulis,
                        unsigned num) {
    pauli_tensor_or_gadget((cx*)raw, n, targets, paulis, num, 0.0, 1.0);
}

/* distributed_statevector.hpp:287-292: exp(+i theta P) = cos(theta) + i sin(theta) P */
void orc_sv_pauliGadget(orc_real* raw, unsigned n, const unsigned* targets, const unsigned* paulis,
                        unsigned num, double theta) {
    pauli_tensor_or_gadget((cx*)raw, n, targets, paulis, num, cos(theta), I * sin(theta));
}

/* local_statevector.hpp:138-153: e^{+i theta} at even parity, e^{-i theta} at odd */
void orc_sv_phaseGadget(orc_real* raw, unsigned n, const unsigned* targets, unsigned num,
                        double theta) {
    cx* amps = (cx*)raw;
    idx_t mask = mask_of(targets, num);
    cx facs[2] = { cos(theta) + I * sin(theta), cos(theta) - I * sin(theta) };
    idx_t numAmps = 1ULL << n;
    #pragma omp parallel for
    for (idx_t j = 0; j < numAmps; j++)
        amps[j] *= facs[parity64(j & mask)];
}

/* ------------------------------------------------------------------------ */
/* density matrix (reference src/distributed_densitymatrix.hpp, 1 rank)      */
/* ------------------------------------------------------------------------ */

static void shifted(unsigned* out, const unsigned* in, unsigned num, unsigned by) {
    for (unsigned i = 0; i < num; i++) out[i] = in[i] + by;
}

static int odd_num_y(const unsigned* paulis, unsigned num) {
    int odd = 0;
    for (unsigned i = 0; i < num; i++) if (paulis[i] == 2) odd = !odd;   /* misc.hpp:47-55 */
    return odd;
}

/* distributed_densitymatrix.hpp:15-25 */
void orc_dm_manyTargGate(orc_real* amps, unsigned N, const unsigned* targets, unsigned k,
                         const orc_real* graw) {
    orc_sv_manyTargGate(amps, 2 * N, targets, k, graw);
    idx_t dim = 1ULL << k;
    orc_real* conjg = (orc_real*)malloc(dim * dim * 2 * sizeof(orc_real));
    for (idx_t i = 0; i < dim * dim; i++) { conjg[2 * i] = graw[2 * i]; conjg[2 * i + 1] = -graw[2 * i + 1]; }
    unsigned t2[64];
    shifted(t2, targets, k, N);
    orc_sv_manyTargGate(amps, 2 * N, t2, k, conjg);
    free(conjg);
}

/* distributed_densitymatrix.hpp:28-33 */
void orc_dm_swapGate(orc_real* amps, unsigned N, unsigned qb1, unsigned qb2) {
    orc_sv_swapGate(amps, 2 * N, qb1, qb2);
    orc_sv_swapGate(amps, 2 * N, qb1 + N, qb2 + N);
}

/* distributed_densitymatrix.hpp:36-50 */
void orc_dm_pauliTensor(orc_real* raw, unsigned N, const unsigned* targets, const unsigned* paulis,
                        unsigned num) {
    orc_sv_pauliTensor(raw, 2 * N, targets, paulis, num);
    unsigned t2[64];
    shifted(t2, targets, num, N);
    orc_sv_pauliTensor(raw, 2 * N, t2, paulis, num);
    if (odd_num_y(paulis, num)) {
        idx_t total = 2ULL << (2 * N);
        #pragma omp parallel for
        for (idx_t j = 0; j < total; j++) raw[j] *= -1;
    }
}

/* distributed_densitymatrix.hpp:53-64 */
void orc_dm_pauliGadget(orc_real* raw, unsigned N, const unsigned* targets, const unsigned* paulis,
                        unsigned num, double theta) {
    orc_sv_pauliGadget(raw, 2 * N, targets, paulis, num, theta);
    unsigned t2[64];
    shifted(t2, targets, num, N);
    if (!odd_num_y(paulis, num)) theta *= -1;
    orc_sv_pauliGadget(raw, 2 * N, t2, paulis, num, theta);
}

/* distributed_densitymatrix.hpp:67-76 */
void orc_dm_phaseGadget(orc_real* raw, unsigned N, const unsigned* targets, unsigned num,
                        double theta) {
    orc_sv_phaseGadget(raw, 2 * N, targets, num, theta);
    unsigned t2[64];
    shifted(t2, targets, num, N);
    orc_sv_phaseGadget(raw, 2 * N, t2, num, -theta);
}

/* distributed_densitymatrix.hpp:79-89 with misc.hpp:58-81 (superoperator
 * sum_K conj(K) (x) K, row r = i*d + k, column c = j*d + l) */
void orc_dm_krausMap(orc_real* raw, unsigned N, const orc_real* kraus, unsigned numOps,
                     const unsigned* targets, unsigned k) {
    idx_t d = 1ULL << k, D = d * d;
    cx* sup = (cx*)calloc(D * D, sizeof(cx));
    for (unsigned o = 0; o < numOps; o++) {
        const cx* K = (const cx*)kraus + (idx_t)o * d * d;
        for (idx_t i = 0; i < d; i++)
            for (idx_t j = 0; j < d; j++)
                for (idx_t kk = 0; kk < d; kk++)
                    for (idx_t l = 0; l < d; l++)
                        sup[(i * d + kk) * D + (j * d + l)] += conj(K[i * d + j]) * K[kk * d + l];
    }
    unsigned ext[128];
    for (unsigned i = 0; i < k; i++) { ext[i] = targets[i]; ext[i + k] = targets[i] + N; }
    orc_sv_manyTargGate(raw, 2 * N, ext, 2 * k, (const orc_real*)sup);
    free(sup);
}

/* local_densitymatrix.hpp:12-42 (1-rank branch :28-41) */
void orc_dm_oneQubitDephasing(orc_real* raw, unsigned N, unsigned qb, double prob) {
    cx* amps = (cx*)raw;
    cx fac = 1 - 2 * prob;
    unsigned alt = qb + N;
    idx_t numIts = (1ULL << (2 * N)) / 4;
    #pragma omp parallel for
    for (idx_t k = 0; k < numIts; k++) {
        idx_t j01 = ins_bit(ins_bit(k, qb, 1), alt, 0);
        amps[j01] *= fac;
        idx_t j10 = ins_bit(ins_bit(k, qb, 0), alt, 1);
        amps[j10] *= fac;
    }
}

/* local_densitymatrix.hpp:45-60 */
void orc_dm_twoQubitDephasing(orc_real* raw, unsigned N, unsigned qb1, unsigned qb2, double prob) {
    cx* amps = (cx*)raw;
    unsigned alt1 = qb1 + N, alt2 = qb2 + N;
    cx term = -4 * prob / 3;
    idx_t num = 1ULL << (2 * N);
    #pragma omp parallel for
    for (idx_t j = 0; j < num; j++) {
        unsigned b1 = (unsigned)(bit_of(j, qb1) ^ bit_of(j, alt1));
        unsigned b2 = (unsigned)(bit_of(j, qb2) ^ bit_of(j, alt2));
        cx flag = (double)(b1 | b2);
        amps[j] *= flag * term + 1.0;
    }
}

/* local_densitymatrix.hpp:63-81 with the constants of distributed_densitymatrix.hpp:106-108 */
void orc_dm_oneQubitDepolarising(orc_real* raw, unsigned N, unsigned qb, double prob) {
    cx* amps = (cx*)raw;
    cx c1 = 2 * prob / 3, c2 = 1 - 2 * prob / 3, c3 = 1 - 4 * prob / 3;
    unsigned alt = qb + N;
    idx_t numIts = (1ULL << (2 * N)) / 4;
    #pragma omp parallel for
    for (idx_t k = 0; k < numIts; k++) {
        idx_t j00 = ins_bit(ins_bit(k, qb, 0), alt, 0);
        idx_t j01 = j00 ^ (1ULL << qb);
        idx_t j10 = j00 ^ (1ULL << alt);
        idx_t j11 = j01 ^ (1ULL << alt);
        cx a00 = amps[j00];
        amps[j00] = c2 * a00 + c1 * amps[j11];
        amps[j01] *= c3;
        amps[j10] *= c3;
        amps[j11] = c1 * a00 + c2 * amps[j11];
    }
}

/* local_densitymatrix.hpp:83-108 loop structure; c1 chosen by the caller */
static void two_qubit_depol_loops(cx* amps, unsigned N, unsigned qb1, unsigned qb2, cx c1, cx c2, cx c3) {
    if (qb1 > qb2) { unsigned s = qb1; qb1 = qb2; qb2 = s; }    /* distributed_densitymatrix.hpp:244-245 */
    unsigned q0 = qb1, q1 = qb2, q2 = qb1 + N, q3 = qb2 + N;
    idx_t num = 1ULL << (2 * N);
    #pragma omp parallel for
    for (idx_t j = 0; j < num; j++) {
        unsigned f1 = !(bit_of(j, q0) ^ bit_of(j, q2));
        unsigned f2 = !(bit_of(j, q1) ^ bit_of(j, q3));
        cx fac = 1.0 + c3 * (double)(!(f1 & f2));
        amps[j] *= fac;
    }
    idx_t numIts = num / 16;
    unsigned sorted[4] = { q0, q1, q2, q3 };
    sort_asc(sorted, 4);
    #pragma omp parallel for
    for (idx_t k = 0; k < numIts; k++) {
        idx_t j0000 = ins_bits(k, sorted, 4, 0);
        idx_t j0101 = j0000 ^ (1ULL << q2) ^ (1ULL << q0);
        idx_t j1010 = j0000 ^ (1ULL << q3) ^ (1ULL << q1);
        idx_t j1111 = j0101 ^ (1ULL << q3) ^ (1ULL << q1);
        cx term = amps[j0000] + amps[j0101] + amps[j1010] + amps[j1111];
        amps[j0000] = c1 * amps[j0000] + c2 * term;
        amps[j0101] = c1 * amps[j0101] + c2 * term;
        amps[j1010] = c1 * amps[j1010] + c2 * term;
        amps[j1111] = c1 * amps[j1111] + c2 * term;
    }
}

/* D3 contract: textbook channel.  Off-block-diagonal entries scale by
 * 1-16p/15 (c3 of distributed_densitymatrix.hpp:249); block-diagonal entries
 * become (1-16p/15) rho_rr + (4p/15) sum_s rho_ss. */
void orc_dm_twoQubitDepolarising(orc_real* raw, unsigned N, unsigned qb1, unsigned qb2, double prob) {
    two_qubit_depol_loops((cx*)raw, N, qb1, qb2, 1 - 16 * prob / 15, 4 * prob / 15, -16 * prob / 15);
}

/* as shipped: distributed_densitymatrix.hpp:247-249 constants, c1 = 1-4p/5 */
void orc_dm_twoQubitDepolarising_asShipped(orc_real* raw, unsigned N, unsigned qb1, unsigned qb2,
                                           double prob) {
    two_qubit_depol_loops((cx*)raw, N, qb1, qb2, 1 - 4 * prob / 5, 4 * prob / 15, -16 * prob / 15);
}

/* local_densitymatrix.hpp:111-131 */
void orc_dm_damping(orc_real* raw, unsigned N, unsigned qb, double prob) {
    cx* amps = (cx*)raw;
    unsigned alt = qb + N;
    cx c1 = sqrt(1 - prob), c2 = 1 - prob;
    idx_t numIts = (1ULL << (2 * N)) / 4;
    #pragma omp parallel for
    for (idx_t k = 0; k < numIts; k++) {
        idx_t j00 = ins_bit(ins_bit(k, qb, 0), alt, 0);
        idx_t j01 = j00 ^ (1ULL << qb);
        idx_t j10 = j00 ^ (1ULL << alt);
        idx_t j11 = j01 ^ (1ULL << alt);
        amps[j00] += prob * amps[j11];
        amps[j01] *= c1;
        amps[j10] *= c1;
        amps[j11] *= c2;
    }
}

/* distributed_densitymatrix.hpp:322-344 with misc.hpp:16-38: the element of
 * the Pauli tensor at flat index i takes sigma[code][bit(i,q+N)][bit(i,q)] */
static const cx PAULI_MATS[4][2][2] = {
    { { 1, 0 }, { 0, 1 } },
    { { 0, 1 }, { 1, 0 } },
    { { 0, -I }, { I, 0 } },
    { { 1, 0 }, { 0, -1 } },
};

void orc_dm_expecPauliString(const orc_real* raw, unsigned N, const double* coeffs, unsigned numTerms,
                             const unsigned* allPaulis, orc_real* out) {
    const cx* amps = (const cx*)raw;
    idx_t num = 1ULL << (2 * N);
    double vr = 0, vi = 0;
    #pragma omp parallel for reduction(+:vr, vi)
    for (idx_t j = 0; j < num; j++) {
        cx term = 0;
        for (unsigned t = 0; t < numTerms; t++) {
            const unsigned* codes = allPaulis + (idx_t)t * N;
            cx elem = 1;
            for (unsigned q = 0; q < N; q++) {
                unsigned colBit = (unsigned)bit_of(j, q);
                unsigned rowBit = (unsigned)bit_of(j, q + N);
                elem *= PAULI_MATS[codes[q]][rowBit][colBit];
            }
            term += elem * coeffs[t];
        }
        cx v = term * amps[j];
        vr += creal(v);
        vi += cimag(v);
    }
    out[0] = vr;
    out[1] = vi;
}

/* distributed_densitymatrix.hpp:347-370 (1-rank branch) and
 * local_densitymatrix.hpp:134-164 with the D1-corrected setBits */
void orc_dm_partialTrace(const orc_real* inRaw, unsigned N, const unsigned* targets, unsigned k,
                         orc_real* outRaw) {
    const cx* in = (const cx*)inRaw;
    cx* out = (cx*)outRaw;
    unsigned targs[64], pairs[64], all[128];
    for (unsigned i = 0; i < k; i++) targs[i] = targets[i];
    sort_asc(targs, k);
    for (unsigned i = 0; i < k; i++) { pairs[i] = targs[i] + N; all[i] = targs[i]; all[i + k] = pairs[i]; }
    sort_asc(all, 2 * k);
    idx_t numOut = 1ULL << (2 * (N - k));
    idx_t numTraced = 1ULL << k;
    #pragma omp parallel for
    for (idx_t l = 0; l < numOut; l++) {
        cx acc = 0;
        idx_t lMask = ins_bits(l, all, 2 * k, 0);
        for (idx_t kk = 0; kk < numTraced; kk++) {
            idx_t i = put_bits(lMask, targs, k, kk);
            i = put_bits(i, pairs, k, kk);
            acc += in[i];
        }
        out[l] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* extensions (no reference counterpart)                                     */
/* ------------------------------------------------------------------------ */

/* P|b>[j0] = powI (-1)^{par(j1 & maskYZ)} b[j1], j1 = j0 ^ maskXY: the
 * pauliTensor map of local_statevector.hpp:129-131 with thisAmpFac = 0. */
static void pauli_masks(const unsigned* codes, unsigned n, idx_t* mxy, idx_t* myz, cx* powI) {
    *mxy = 0; *myz = 0; *powI = 1;
    for (unsigned q = 0; q < n; q++) {
        unsigned c = codes[q];
        if (c == 1 || c == 2) *mxy |= 1ULL << q;
        if (c == 2 || c == 3) *myz |= 1ULL << q;
        if (c == 2) *powI *= I;
    }
}

void orc_sv_applyPauliSum(const orc_real* inRaw, orc_real* outRaw, unsigned n, const double* coeffs,
                          unsigned numTerms, const unsigned* paulis) {
    const cx* in = (const cx*)inRaw;
    cx* out = (cx*)outRaw;
    idx_t num = 1ULL << n;
    idx_t* mxy = (idx_t*)malloc(numTerms * sizeof(idx_t));
    idx_t* myz = (idx_t*)malloc(numTerms * sizeof(idx_t));
    cx* fac = (cx*)malloc(numTerms * sizeof(cx));
    for (unsigned t = 0; t < numTerms; t++) {
        cx p;
        pauli_masks(paulis + (idx_t)t * n, n, &mxy[t], &myz[t], &p);
        fac[t] = coeffs[t] * p;
    }
    #pragma omp parallel for
    for (idx_t j = 0; j < num; j++) {
        cx acc = 0;
        for (unsigned t = 0; t < numTerms; t++) {
            idx_t j1 = j ^ mxy[t];
            double s = 1.0 - 2.0 * parity64(j1 & myz[t]);
            acc += fac[t] * s * in[j1];
        }
        out[j] = acc;
    }
    free(mxy); free(myz); free(fac);
}

void orc_sv_innerProduct(const orc_real* aRaw, const orc_real* bRaw, unsigned n, orc_real* out) {
    const cx* a = (const cx*)aRaw;
    const cx* b = (const cx*)bRaw;
    idx_t num = 1ULL << n;
    double vr = 0, vi = 0;
    #pragma omp parallel for reduction(+:vr, vi)
    for (idx_t j = 0; j < num; j++) {
        cx v = conj(a[j]) * b[j];
        vr += creal(v);
        vi += cimag(v);
    }
    out[0] = vr;
    out[1] = vi;
}

void orc_sv_pauliMatrixElement(const orc_real* aRaw, const orc_real* bRaw, unsigned n,
                               const unsigned* targets, const unsigned* paulis, unsigned num,
                               orc_real* out) {
    const cx* a = (const cx*)aRaw;
    const cx* b = (const cx*)bRaw;
    idx_t mxy = 0, myz = 0;
    cx powI = 1;
    for (unsigned i = 0; i < num; i++) {
        if (paulis[i] == 1 || paulis[i] == 2) mxy ^= 1ULL << targets[i];
        if (paulis[i] == 2 || paulis[i] == 3) myz ^= 1ULL << targets[i];
        if (paulis[i] == 2) powI *= I;
    }
    idx_t numAmps = 1ULL << n;
    double vr = 0, vi = 0;
    #pragma omp parallel for reduction(+:vr, vi)
    for (idx_t j = 0; j < numAmps; j++) {
        idx_t j1 = j ^ mxy;
        double s = 1.0 - 2.0 * parity64(j1 & myz);
        cx v = conj(a[j]) * (powI * s * b[j1]);
        vr += creal(v);
        vi += cimag(v);
    }
    out[0] = vr;
    out[1] = vi;
}
