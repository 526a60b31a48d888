// dense.h -- tall-skinny dense kernels of the block Lanczos eigensolver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>

// out = alpha * V^T W + beta * out   (V: n x P, W: n x Q <= 16; out ld ldo)
void ts_gemm_tn(int n, int P, int Q, const double *V, int ldv, const double *W, int ldw, double *out, int ldo,
                double alpha, double beta, double *part, hipStream_t s);
int64_t ts_gemm_tn_part_size(int n, int P, int Q);
// out = V^T W and out2 += V^T W (Q <= 16)
void ts_gemm_tn_add(int n, int P, int Q, const double *V, int ldv, const double *W, int ldw, double *out, int ldo,
                    double *out2, int ldo2, double *part, hipStream_t s);
// Y = alpha * V H + beta * Y   (Q <= 16)
void ts_gemm_nn(int n, int P, int Q, const double *V, int ldv, const double *H, int ldh, double *Y, int ldy,
                double alpha, double beta, hipStream_t s);
// as ts_gemm_nn, and G = Y^T Y (Q x Q, ld ldg) of the result; part holds
// ts_gemm_tn_part_size(n, 16, 16) doubles
void ts_gemm_nn_gram(int n, int P, int Q, const double *V, int ldv, const double *H, int ldh, double *Y, int ldy,
                     double alpha, double beta, double *part, double *G, int ldg, hipStream_t s);
// Y -= V H, G = Y^T Y (Q x Q, ld ldg), C = V^T Y (P x Q, ld ldc) and
// Cadd += C, in one pass over V for P <= 512 a multiple of 16 (Q <= 16); part holds
// max(ts_gemm_nn_tn_part_size, ts_gemm_tn_part_size) doubles
void ts_gemm_nn_gram_tn(int n, int P, int Q, const double *V, int ldv, const double *H, int ldh, double *Y, int ldy,
                        double *part, double *G, int ldg, double *C, int ldc, double *Cadd, int ldca, hipStream_t s);
int64_t ts_gemm_nn_tn_part_size(int n, int P, int Q);
void small_chol_upper(int b, double *G, int ldg, int *err, hipStream_t s);

// as small_chol_upper, but a pivot <= tol2 * max diag(g0) flags block blk as
// rank-deficient (err[0] |= 4, err[1] = min(blk + 1))
void small_chol_upper_checked(int b, double *G, int ldg, const double *g0, int ldg0, double tol2, int blk, int *err,
                              hipStream_t s);
// V[:, 0:cols] = seeded uniform values in [-1, 1)
void hash_fill(int n, int cols, double *V, int ldv, uint64_t seed, hipStream_t s);
void trsm_right_upper(int n, int b, double *W, int ldw, const double *R, int ldr, hipStream_t s);
// as trsm_right_upper, and Rout = R R1 (R1 upper b x b)
void trsm_right_upper_mul(int n, int b, double *W, int ldw, const double *R, int ldr, const double *R1, int ld1,
                          double *Rout, int ldo, hipStream_t s);
// as trsm_right_upper, and G = W^T W (b x b, ld ldg) of the result
void trsm_right_upper_gram(int n, int b, double *W, int ldw, const double *R, int ldr, double *part, double *G,
                           int ldg, hipStream_t s);
// The two Cholesky-QR passes without separate factorisation launches: every
// workgroup factors the b x b Gram block G itself (R = chol(G), the
// arithmetic of small_chol_upper(_checked), same flags in err).
// chol_trsm_right_upper_gram: W = W R^-1, G2 = W^T W, R stored to Rsave (ld 16);
// g0 non-null checks for dependent directions as small_chol_upper_checked.
void chol_trsm_right_upper_gram(int n, int b, double *W, int ldw, const double *G, int ldg, const double *g0,
                                int ldg0, double tol2, int blk, int *err, double *Rsave, double *part, double *G2,
                                int ldg2, hipStream_t s);
// W = W R^-1 and, with R1 non-null, Rout = R R1
void chol_trsm_right_upper(int n, int b, double *W, int ldw, const double *G, int ldg, int *err, const double *R1,
                           int ld1, double *Rout, int ldo, hipStream_t s);
void upper_mul(int b, const double *R2, int ld2, const double *R1, int ld1, double *out, int ldo, hipStream_t s);
void start_block(int n, int b, double *V, int ldv, int seed, hipStream_t s);
// (max, sum) over the shells (CQUAD4, CTRIA3) of the mean |translation|^2
// at their nodes, for mode column k
void elem_energy_share(int n_q4, const int32_t *q4, int n_t3, const int32_t *t3, int ndof, const double *modes,
                       int k, double *scratch, double *out2, hipStream_t s);
void normalize_maxabs(int64_t n, double *x, double *scratch, hipStream_t s);
void symmetrize_copy(int m, int ld, const double *T, double *A, hipStream_t s);
void dot_det(int64_t n, const double *x, const double *y, double *scratch, double *out, hipStream_t s);
// g[6v + c] = sum of the element vectors (CQUAD4 24, CTRIA3 18, CBAR 12 per
// element) of the elements at node v, in the block gather's element order;
// zero on fixed DOFs
void node_gather_vec(int n_nodes, int n_q4, int n_t3, const int32_t *diag_blk, const int32_t *cptr,
                     const int32_t *cidx, const double *eg_q4, const double *eg_t3, const double *eg_bar,
                     const uint8_t *free_mask, double *g, hipStream_t s);
void extract_kdiag(int n_nodes, const int32_t *diag_blk, const double *kv, double *out, hipStream_t s);
// dst[i + j*ldd] += src[i + j*lds]
void mat_add(int rows, int cols, const double *src, int lds, double *dst, int ldd, hipStream_t s);
// The k eigenpairs of largest |theta| (algebraic != 0: of largest theta) of a
// symmetric m x m matrix A (column-major, may be overwritten), one workgroup;
// m <= SYMEIG_MAX_M, k <= SYMEIG_MAX_K.  theta[t] in descending |theta|
// (descending) order, Zs (m x k, ld m)
// the orthonormal eigenvectors; work holds work_n >= symeig_work_size(m, k)
// doubles; *info = 0 on success.  bw > 0 declares A banded with half-bandwidth bw
// (block Lanczos without restarts): for 4 <= bw <= SYMEIG_BAND_MAX_W the
// pairs then come from a band reduction (the band in LDS up to
// SYMEIG_BAND_MAX_M, in the workspace above), which checks the claim
// (off-band entries within 1e-12 max|A|) and otherwise leaves the matrix to
// the dense path.
constexpr int SYMEIG_MAX_M = 1024;
constexpr int SYMEIG_MAX_K = 64;
constexpr int SYMEIG_BAND_MAX_M = 384;
constexpr int SYMEIG_BAND_MAX_W = 16;
int64_t symeig_work_size(int m, int k);
void sym_eig_top(int m, double *A, int k, double *theta, double *Zs, double *work, int64_t work_n, int *info,
                 hipStream_t s, int bw = 0, int algebraic = 0);

// ---------------------------------------------------------------------------
// Segmented forms for a batch of independent problems stacked by rows (each
// member's free DOFs in elimination order).  Every operation acts on each
// active segment alone; per-segment small matrices are strided (base + s *
// stride).  Chunks never cross a segment and start at multiples of their
// size from the segment start, so a segment reproduces the single-problem
// arithmetic.
// ---------------------------------------------------------------------------
struct SegTable {
    int n_seg = 0, n_ch = 0, n_chs = 0;
    const int32_t *seg_off = nullptr;                                         // n_seg + 1
    const int32_t *ch_r0 = nullptr, *ch_seg = nullptr, *ch_ptr = nullptr;     // 512-row chunks, ptr n_seg + 1
    const int32_t *chs_r0 = nullptr, *chs_seg = nullptr, *chs_ptr = nullptr;  // 256-row chunks
    const uint8_t *active = nullptr;                                          // n_seg
};
constexpr int SEG_CH = 512, SEG_CHS = 256;
void seg_gemm_tn(const SegTable &T, int P, int Q, const double *V, int ldv, const double *W, int ldw, double *out,
                 int64_t so, int ldo, double alpha, double beta, double *part, hipStream_t s);
int64_t seg_gemm_tn_part_size(int n_ch, int n_chs, int P, int Q);
// out = V^T W and out2 += V^T W per segment
void seg_gemm_tn_add(const SegTable &T, int P, int Q, const double *V, int ldv, const double *W, int ldw, double *out,
                     int64_t so, int ldo, double *out2, int64_t so2, int ldo2, double *part, hipStream_t s);
// Q <= 16
void seg_gemm_nn(const SegTable &T, int P, int Q, const double *V, int ldv, const double *H, int64_t sh, int ldh,
                 double *Y, int ldy, double alpha, double beta, hipStream_t s);
// as seg_gemm_nn, and G = Y^T Y (Q x Q per segment, stride sg) of the result
void seg_gemm_nn_gram(const SegTable &T, int P, int Q, const double *V, int ldv, const double *H, int64_t sh, int ldh,
                      double *Y, int ldy, double alpha, double beta, double *part, double *G, int64_t sg, int ldg,
                      hipStream_t s);
// Y -= V H, G = Y^T Y, C = V^T Y and Cadd += C per segment in one pass over
// V (ts_gemm_nn_gram_tn per segment); part holds
// max(seg_gemm_nn_tn_part_size, seg_gemm_tn_part_size) doubles
void seg_gemm_nn_gram_tn(const SegTable &T, int P, int Q, const double *V, int ldv, const double *H, int64_t sh,
                         int ldh, double *Y, int ldy, double *part, double *G, int64_t sg, int ldg, double *C,
                         int64_t sc, int ldc, double *Cadd, int64_t sca, int ldca, hipStream_t s);
int64_t seg_gemm_nn_tn_part_size(int n_ch, int P, int Q);
// err: two ints per segment, as small_chol_upper(_checked) sets them
void seg_chol_upper(const SegTable &T, int b, double *G, int64_t sg, int ldg, const double *g0, int64_t sg0,
                    int ldg0, double tol2, int *err, hipStream_t s);
void seg_trsm_right_upper(const SegTable &T, int b, double *W, int ldw, const double *R, int64_t sr, int ldr,
                          hipStream_t s);
// as seg_trsm_right_upper, and Rout = R R1 per segment
void seg_trsm_right_upper_mul(const SegTable &T, int b, double *W, int ldw, const double *R, int64_t sr, int ldr,
                              const double *R1, int64_t s1, int ld1, double *Rout, int64_t so, int ldo, hipStream_t s);
// as seg_trsm_right_upper, and G = W^T W (b x b per segment) of the result
void seg_trsm_right_upper_gram(const SegTable &T, int b, double *W, int ldw, const double *R, int64_t sr, int ldr,
                               double *part, double *G, int64_t sg, int ldg, hipStream_t s);
void seg_upper_mul(const SegTable &T, int b, const double *R2, int64_t s2, int ld2, const double *R1, int64_t s1,
                   int ld1, double *out, int64_t so, int ldo, hipStream_t s);
// start block of every segment (row index local to the segment)
void seg_start_block(const SegTable &T, int b, double *V, int ldv, int seed, hipStream_t s);
// V[rows of the active segments, 0:cols] = v
void seg_fill(const SegTable &T, int cols, double *V, int ldv, double v, hipStream_t s);
void seg_symmetrize_copy(const SegTable &T, int m, int ldT, const double *Tm, int64_t sT, double *A, int64_t sA,
                         hipStream_t s);
void seg_mat_add(const SegTable &T, int rows, int cols, const double *src, int64_t ss, int lds, double *dst,
                 int64_t sd, int ldd, hipStream_t s);
// dst_s = src_s (rows x cols) for the active segments
void seg_copy(const SegTable &T, int rows, int cols, const double *src, int64_t ss, int lds, double *dst, int64_t sd,
              int ldd, hipStream_t s);
// per (member, column) pair: (max, sum) over the member's shells (element
// ranges q4_off / t3_off) of the mean |translation|^2 at their nodes
void seg_energy_share(int n_pairs, const int32_t *pairs, const int32_t *q4_off, const int32_t *q4,
                      const int32_t *t3_off, const int32_t *t3, int64_t ldx, const double *X, double *out2,
                      hipStream_t s);
// per (member, column) pair: X[dof_off[m]:dof_off[m+1], col] / its max |x|, to dst + dst_off[pair]
void seg_normalize(int n_pairs, const int32_t *pairs, const int64_t *dof_off, const int64_t *dst_off, int64_t ldx,
                   const double *X, double *dst, hipStream_t s);
// sym_eig_top for the listed problems p (A + p sA, theta + p sT, Zs + p sZ,
// work + p sW with sW doubles each, info + p), one workgroup each, all of size m
void sym_eig_top_batched(int m, double *A, int64_t sA, int k, double *theta, int64_t sT, double *Zs, int64_t sZ,
                         double *work, int64_t sW, int *info, const int *list, int nlist, hipStream_t s, int bw = 0);
