// Internal kernel interface of libqrochet_hip.so (gfx950).
//
// Device data layout of one MPS site tensor A[l, o, r]: complex FP64, l fastest,
// then o, then r (column-major (l, o, r)).  The same memory is the (χl·d) × χr
// matrix with rows (l, o) and the χl × (d·χr) matrix with columns (o, r), so the
// two-site contraction and both SVD factors need no permutation.
//
// Batched launches take an array of item descriptors and a tile map (int2:
// item, tile-within-item), both in device memory, so one launch serves items
// of different shapes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qh {

using cplx = double2;

// ---------------------------------------------------------------------------
// θ build (two-site contraction + gate), one item per bond update
// ---------------------------------------------------------------------------
struct ThetaItem {
    const cplx* al;      // A_l  (χl, d, χm)
    const cplx* ar;      // A_r  (χm, d, χr)
    const double* lam;   // Λ on the bond (χm) or null
    const double* rs;    // left scale on l (χl) or null  (Vidal mode: Λ_{i-1})
    const double* cs;    // right scale on r (χr) or null (Vidal mode: Λ_{i+1})
    const cplx* gate;    // d^4, (out_l, out_r, in_l, in_r) column-major
    cplx* x;             // output: θ (m × n, ld m) or θ^H (n × m, ld n)
    int chil, chim, chir, d;
    int transposed;      // 0: X = θ,  1: X = θ^H
    int tiles_l, tiles_r;
};
// FP64 vector-FMA build; d in {2, 3, 4}.  Tiles cover theta_tile_extent(d)
// values of l and of r.
void launch_theta(const ThetaItem* d_items, const int2* d_tiles, int ntiles, int d, hipStream_t s);
int theta_tile_extent(int d);
// Reduced-precision build on MFMA (precision = QH_PREC_FP32 or QH_PREC_BF16X3), FP64
// gate mix and output; tiles cover 32 values of l and of r.
void launch_theta_mfma(const ThetaItem* d_items, const int2* d_tiles, int ntiles, int d, int precision,
                       hipStream_t s);
constexpr int THETA_MFMA_TILE = 32;

// ---------------------------------------------------------------------------
// batched complex GEMM:  C = alpha · op(A) · diag(dk) · op(B) + beta · C
// then C rows scaled by rowscale, columns by colscale (nullable)
// ---------------------------------------------------------------------------
struct GemmItem {
    const cplx* a;
    const cplx* b;
    cplx* c;
    const double* dk;        // K scaling, index ((k + dk_off) / dk_div) % dk_mod (null = none)
    const double* rowscale;  // M
    const double* colscale;  // N
    cplx alpha, beta;
    int m, n, k, lda, ldb, ldc;
    int opa, opb;            // GEMM_N, GEMM_C (conjugate transpose), GEMM_T (transpose), GEMM_CONJ
    int dk_mod, dk_div;
    int tiles_m, tiles_n;
    int dk_off;              // K offset of the scaling index (a K-slice of a larger product)
};
enum { GEMM_N = 0, GEMM_C = 1, GEMM_T = 2, GEMM_CONJ = 3 };
// element (r, c) of op(X) where X is stored column-major with leading dimension ld
__device__ __forceinline__ bool op_plain(int op) { return op == GEMM_N || op == GEMM_CONJ; }
__device__ __forceinline__ cplx load_op(const cplx* x, int ld, int r, int c, int op) {
    if (op_plain(op)) {
        const cplx v = x[r + (size_t)ld * c];
        return op == GEMM_CONJ ? make_double2(v.x, -v.y) : v;
    }
    const cplx v = x[c + (size_t)ld * r];
    return op == GEMM_C ? make_double2(v.x, -v.y) : v;
}
// FP64 on the vector ALUs
void launch_gemm(const GemmItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// split-K epilogue: C = alpha (Σ_s W_s) rowscale colscale (+ beta C), the slices W_s
// (m × n each, ld m, consecutive) added in order s = 0, 1, ...
struct GemmReduceItem {
    const cplx* w;
    int slices;
    cplx* c;
    int m, n, ldc;
    cplx alpha, beta;
    const double* rowscale;
    const double* colscale;
};
// tile map: (item, block of 256 entries of the m × n output)
void launch_gemm_reduce(const GemmReduceItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// the same kernel for one GEMM passed by value (tiles and divisors filled in here)
void launch_gemm_arg(GemmItem g, hipStream_t s);
// FP32 or bf16x3 products on MFMA with FP32 accumulation (precision = QH_PREC_FP32 /
// QH_PREC_BF16X3), FP64 epilogue; same tiles as launch_gemm
void launch_gemm_mfma(const GemmItem* d_items, const int2* d_tiles, int ntiles, int precision, hipStream_t s);
constexpr int GEMM_TM = 64, GEMM_TN = 64, GEMM_TK = 16;

// ---------------------------------------------------------------------------
// batched block one-sided Jacobi SVD
// ---------------------------------------------------------------------------
struct JacItem {
    cplx* x;        // p × q working matrix (ld p), q <= p, columns orthogonalised in place
    cplx* v;        // q × q accumulated right rotations (ld q)
    int p, q;
    int nb;         // column blocks (even >= 2)
    int* rot;       // rotations performed in the current window (device counter)
    int* conv;      // 1 once converged (kernels skip the item)
    int* tail;      // 1 while the item is in the tail passes (the rounds skip it), 2 once it left them
    int* own;       // JAC_OWN_INTS ints: rotated block pairs of the item's own sweeps (by
                    // sweep index mod 5), and the tail's bookkeeping (below)
    double* fro2;   // ||X||_F^2 (device scalar, filled before the sweeps)
    double tol;     // relative orthogonality target of this matrix
    double floor_rel2;  // deflation floor relative to ||X||_F^2
};
// JacItem::own: [0, 5) rotated block pairs by own sweep mod 5; the item's round offset
// (its own round = the batch's round - offset, set when it returns from the tail so that
// it resumes its own sweep at its first round); the own sweep at which it entered the
// tail; the first own sweep at which it may enter the tail.
constexpr int JAC_ROUND_OFF = 5, JAC_TAIL_AT = 6, JAC_TAIL_MIN = 7, JAC_OWN_INTS = 8;
// Columns with ||x||^2 <= JAC_FLOOR(p)^2 ||X||_F^2 are below the rounding floor of the
// rotations and are treated as zero (frozen): JAC_FLOOR(p) = 4 sqrt(p) eps.
// ||X||_F^2 per item (tile map: item, chunk of 64k elements)
void launch_frob2(const JacItem* d_items, int count, const int2* d_tiles, int ntiles, double* partial,
                  const int* first, hipStream_t s);
constexpr int FROB_CHUNK = 65536;
// block = columns per block (8 or 16); a workgroup handles 2·block columns.
// round: the tournament round, plus *round_base when round_base is not null
// tail_switch > 0 (per mille): at the start of each of its own sweeps after
// the sorting ones, an item whose previous sweep rotated at most that share of its block
// pairs stops (tail = 1) and waits for the tail passes (JAC_CHECK / JAC_PAIRS launches)
// vq / vhdr / vmap (block width 16 only; vq may be null): for the nv pairs of matrices
// with at least JAC_DEFER_MIN_Q columns, V_IJ <- V_IJ Q is left to the V workgroups of the
// next launch.  vmap[t] (t < ntiles) is pair t's V index v or -1, vmap[ntiles + v] the
// pair of V index v.  Pair t stores Q (2·16 × 2·16) and its blocks in slot
// (g & 1)·nv + v (g the global round; 2·nv slots of Q, 4 ints of header each: flag, bi, bj),
// and the launch carries nv V workgroups that apply the previous round's slots.  A window
// of rounds ends with launch_jacobi_vonly (round = the window's end), which applies the
// last round's.  The headers start at zero.
constexpr int JAC_DEFER_MIN_Q = 256;
void launch_jacobi_round(const JacItem* d_items, const int2* d_tiles, int ntiles, int round,
                         const int* round_base, int block, int sort_sweeps, int debug, int max_inner, int tail_switch,
                         cplx* vq, int* vhdr, const int* vmap, int nv, hipStream_t s);
void launch_jacobi_vonly(const JacItem* d_items, const int2* d_tiles, int ntiles, int round, const int* round_base,
                         cplx* vq, int* vhdr, const int* vmap, int nv, hipStream_t s);
// Pair k of round `round` of the circle-method tournament over nidx (even) blocks.
__host__ __device__ inline void rr_pair(int round, int k, int nidx, int& a, int& b) {
    const int m = nidx - 1;
    if (k == 0) {
        a = round;
        b = m;
    } else {
        a = (round + k) % m;
        b = (round - k + m) % m;
    }
}
// Modes of the round kernels (JAC_CHECK at block width 8 only): a tournament round; the test of every (round,
// pair) of a sweep in one launch, without rotating (tile: item, round · nb/2 + pair;
// flags[flag0[item] + tile.y] = 1 where the pair needs a rotation); explicit block
// pairs (tile: item, bi | bj << 16), which must be disjoint within an item.
enum JacMode { JAC_ROUND = 0, JAC_CHECK = 1, JAC_PAIRS = 2, JAC_VONLY = 3 };
void launch_jacobi_check(const JacItem* d_items, const int2* d_tiles, int ntiles, int* d_flags, const int* d_flag0,
                         int block, hipStream_t s);
void launch_jacobi_pairs(const JacItem* d_items, const int2* d_tiles, int ntiles, int block, int max_inner,
                         hipStream_t s);
// block width 16 or 32: the test of every column pair of whole items (tile: item, ti, tj
// with ti <= tj, 64-column tiles of the upper triangle of 16-column blocks); sets
// flags[first[item] + I·nb + J] = 1 for the failing blocks I <= J (the table must start
// at zero)
void launch_jacobi_check_tiles(const JacItem* d_items, const int4* d_tiles, int ntiles, int* d_flags,
                               const int* d_first, int block, hipStream_t s);
// V <- I, counters cleared (tile map: item, column)
void launch_set_identity_items(const JacItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// converged <- (no rotation in the window, outside the tail); remaining[0] = items still
// rotating, remaining[1] = items with non-finite entries, remaining[3] = items waiting
// for the tail passes (an item whose own sweep ends with the window, at first_round +
// rounds, takes the round kernel's tail decision here); if round_base is not null,
// *round_base += rounds (the next sweep's first round)
void launch_conv_update(const JacItem* d_items, int count, int* remaining, int* round_base, int rounds,
                        int first_round, int sort_sweeps, int tail_switch, hipStream_t s);

// Finalisation: σ_j = ||X[:, j]||, sort descending, truncation rule of
// `truncate!`, optional renormalisation, kept count and discarded weight.
struct SelectItem {
    const cplx* x;
    int p, q;
    double* sigma_sorted;   // q (in: unsorted norms; out: descending)
    int* perm;              // q (source column of each sorted value)
    double* lam_out;        // kept Λ (k'), renormalised if requested
    int* kept;              // device scalar
    double* discarded;      // device scalar
    double threshold;       // already defaulted (1e-16 if only maxdim)
    int maxdim;             // < 0 -> none
    int truncate;           // 0 -> keep all q
    int renormalize;
};
constexpr int SELECT_MAX_Q = 4096;
// norm tiles: (item, group of 4 columns); one select workgroup per item
void launch_select_items(const SelectItem* d_items, const int2* d_norm_tiles, int n_norm_tiles, int count,
                         hipStream_t s);

// Kept columns at or below the deflation floor (sorted positions first..kept-1 with
// σ_j² <= floor_rel2·||X||_F², the Jacobi never rotated them) carry no direction of
// their own: their singular vectors are replaced by an orthonormal completion of
// the columns above the floor, written back into the Jacobi matrix as σ_j z_j
// (z_j a unit vector) so that the gathers' division by σ_j yields z_j; a σ_j below
// 1e-280 is reported as 0 and z_j stored unscaled.  One workgroup per item; items
// without such columns return at once.
struct CompleteItem {
    cplx* x;                 // Jacobi matrix (p × q, ld p): sorted column j is column perm[j]
    int p;
    const int* perm;
    double* sigma_sorted;
    double* lam_out;         // kept Λ as written by select (reset to 0 with σ_j)
    const int* kept;
    const double* fro2;      // ||X||_F² before the sweeps
    double floor_rel2;
};
// random starts of the frozen columns, orthogonalised against the columns above the
// floor (tile map: item, sorted column j, for every j below the item's kept bound)
void launch_complete_project(const CompleteItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// d_counts[2i] = the first frozen sorted position of item i, d_counts[2i + 1] = its kept
void launch_complete_count(const CompleteItem* d_items, int count, int* d_counts, hipStream_t s);
// σ (and Λ) below 1e-280 in the frozen positions set to 0
void launch_complete_zero_tiny(const CompleteItem* d_items, const int* d_counts, int count, hipStream_t s);
void launch_eye_rect(cplx* p, int m, int n, hipStream_t s);

// x[r, perm[j]] = f[r, j] (times σ_j if mul_sigma and σ_j > 0) for j < kept: the
// inverse of a gather (tile map: item, column j)
struct ScatterItem {
    cplx* x;
    int p;
    const int* perm;
    const double* sigma_sorted;
    const int* kept;
    const cplx* f;
    int ldf;
    int mul_sigma;
};
void launch_scatter_cols(const ScatterItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);

// out[r, j] = src[r, perm[j]] * colfac(j) * rowscale[(r / div) % mod] for j < kept
//   colfac = 1/sigma_sorted[j] if divide_sigma (1 if sigma == 0: see CompleteItem) else 1
//   conj_transpose: out is (kept × rows), ld kept, holding the conjugate
struct GatherItem {
    const cplx* src;
    int ld_src;
    int rows;
    const int* perm;
    const double* sigma_sorted;
    int divide_sigma;
    const int* kept;
    const double* rowscale;
    int rowscale_mod, rowscale_div;
    cplx* out;
    int conj_transpose;
    int max_cols;
    int ld_out;     // non-transposed output leading dimension (0: rows)
    const int* row_src = nullptr;  // output row c reads source row row_src[c] (nullptr: c)
};
// tile map: (item, tile) over 32 × 32 tiles, tile = jt * ceil(rows/32) + ct
void launch_gather_items(const GatherItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);

// ---------------------------------------------------------------------------
// batched blocked Householder QR (preconditioner of the Jacobi SVD)
// ---------------------------------------------------------------------------
constexpr int QR_NB = 16;  // panel width
struct QrItem {
    cplx* a;        // m × n (m >= n), factored in place: R in the upper triangle
    int m, n, lda;
    cplx* v;        // explicit reflectors (m × n, ld m): zero above the diagonal, 1 on it
    cplx* t;        // compact-WY triangles, panel k at t + k·QR_NB², QR_NB × QR_NB column-major
};
struct QrApplyItem {
    const cplx* v;  // reflectors of a QrItem (ld m)
    const cplx* t;
    int m;          // rows of the reflectors / target
    int nrefl;      // number of reflectors (n of the factorisation)
    cplx* y;        // target (m × ncols, ld ldy)
    int ldy;
    int ncols;
    int trailing;   // 1: columns from j0 + panel width (factorisation update); 0: all columns
};
// panel j0 of every item (grid = items; items with n <= j0 return at once); every
// item has at most 512 rows
void launch_qr_panel(const QrItem* d_items, int count, int j0, hipStream_t s);
// the same for any height, 512-thread workgroups (up to 1024 rows in registers)
void launch_qr_panel_tall(const QrItem* d_items, int count, int j0, hipStream_t s);
// apply panel j0's block reflector: adjoint = 1 -> Q_panel^H (factorisation), 0 -> Q_panel
// tile map: (item, column tile of qr_tile_cols() columns)
void launch_qr_apply(const QrApplyItem* d_items, const int2* d_tiles, int ntiles, int j0, int adjoint, hipStream_t s);
// factorisation look-ahead: panel j0 + QR_NB of every item (workgroups [0, count),
// at most 512 rows) and panel j0's trailing update on the tiles (item, k >= 1) beyond
// it, in one launch
void launch_qr_step(const QrItem* d_items, int count, const QrApplyItem* d_app, const int2* d_tiles, int ntiles,
                    int j0, hipStream_t s);
// y <- Q y with all block reflectors of each item (tile map: item, tile of QR_TC columns)
void launch_qr_form(const QrApplyItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
int qr_tile_cols();
// dst (n × n) = (upper triangle of src)^H, zero above the diagonal, for a batch
struct UpperHItem {
    const cplx* src;
    int lds, n;
    cplx* dst;
};
// tile map: (item, block of 256 entries of the n × n output)
void launch_upper_h_batch(const UpperHItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// column order of the first preconditioning pass: the columns of x (p × q, ld p) by
// decreasing norm, ties in index order; perm[k] = the column placed k-th, perm_inv its
// inverse, *count = q (the column count a gather of the permuted copy reads)
constexpr int PIVOT_MAXQ = 4096;
struct PivotItem {
    const cplx* x;
    int p, q;
    int* perm;
    int* perm_inv;
    int* count;
};
void launch_pivot_order(const PivotItem* d_items, int count, hipStream_t s);

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// T[l, μ, r] <- Σ_ν Op[μ + D ν] T[l, ν, r], D in {2,3,4,8,9,16}
void launch_apply_op(cplx* t, int left, int dim, int right, const cplx* op, hipStream_t s);
// A <- A · f(λ) along l (mode 0) or r (mode 2); f = pinv ? pinv(λ; atol) : λ
void launch_scale_bond(cplx* a, int chil, int d, int chir, const double* scale, int mode, int pinv,
                       double atol, hipStream_t s);
void launch_pinv(const double* lam, int n, double atol, double* out, hipStream_t s);
// dst (cols × rows) = src^H, src rows × cols (ld rows)
void launch_conj_transpose(const cplx* src, int rows, int cols, cplx* dst, hipStream_t s);
// ||G - I||_F
void launch_gram_dev(const cplx* gram, int n, double* dev_out, hipStream_t s);
// Σ_{b,a} E[b,a] R[a,b] λb[b] λa[a]
void launch_env_close(const cplx* e, const cplx* r, int nb, int na, const double* lb, const double* la, cplx* out,
                      hipStream_t s);
void launch_fill(double* p, size_t n, double v, hipStream_t s);
// a local operator on the middle index of many tensors T[l, μ, r] (dim <= 16)
struct ApplyItem {
    cplx* t;
    int left, dim, right;
    const cplx* op;
};
// tile map: (item, block of 256 (l, r) pairs)
void launch_apply_op_batch(const ApplyItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);
// closing dots Σ_{b,a} E[b, a] R[a, b] λb[b] λa[a] of many environments (one workgroup each)
struct CloseItem {
    const cplx* e;
    const cplx* r;
    int nb, na;
    const double* lb;
    const double* la;
    cplx* out;
};
void launch_env_close_batch(const CloseItem* d_items, int count, hipStream_t s);
// p[e] <- uniform [0, 1) (+ i·uniform [0, 1) if complex_entries), a hash of (seed, key, e)
void launch_rand_uniform(cplx* p, size_t n, unsigned long long seed, unsigned long long key, int complex_entries,
                         hipStream_t s);
// p[e] <- f · p[e]
void launch_scale_real(cplx* p, size_t n, double f, hipStream_t s);
// column j of a rows × cols matrix times 10^(-decades · j / (cols - 1))
void launch_scale_cols_geometric(cplx* p, int rows, int cols, double decades, hipStream_t s);
void launch_eye(cplx* p, int n, hipStream_t s);

}  // namespace qh
