// Gram transfers of the boundary absorb with each double layer applied as its
// two single-layer factors (ket and bra) instead of the dense double layer.
//
// A transfer step maps rows of a tensor over two pair legs (c1, c2) to rows
// over two other pair legs (o1, o2) of the same layer:
//   Y[r, o1, o2] = sum_{c1, c2} X[r, c1, c2] E[c1, c2, o1, o2],
//   E = sum_phi S'(c1k, c2k, phi, o1k, o2k) S'(c1b, c2b, phi, o1b, o2b),
// with every pair leg in the parity-sorted (ket, bra) order of the double
// layers. Per row that is two small products through the single layer S'
// (2 D^6 d multiply-adds per row instead of D^8 for the dense double layer at
// bond dimension D), run in one kernel with the half-contracted row kept in
// LDS. Rows and both ends use the folded storage of the absorb's intermediates
// (the last pair leg keeps one parity half).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <vector>

namespace spp {

constexpr int kKronMaxTuple = 64;   // tuples of two single legs (D <= 8)
// Role assignments (stages) tabulated per (site, label): 0 contracts the
// layer's N and E legs into (W, S), 1 contracts E and S into (W, N), 2
// contracts W and N into (E, S), 3 contracts S and W into (E, N). An absorb
// runs 0 then 1; an absorb in the rotated frame (W', N', E', S') = (E, S, W, N)
// runs 3 then 2, which are its 0 and 1 in that frame.
constexpr int kKronStages = 4;

// Lists of one stage (role assignment) of one (site, label), followed in
// memory by the FP32 coefficients K1 [2][dphi][j1cap][no1] and
// K2 [2][dphi][j2cap][no2]. The inner index phi is taken one value at a time:
// pass phi of stage 1 yields the outputs (o1k, o2k) of that phi, pass phi of
// stage 2 sums them.
struct KronTab {
  int n12[2];                    // tuples (c1, c2) of parity pi
  int no[2];                     // tuples (o1, o2) of parity pi
  int nj2[2][2];                 // stage-2 inner tuples (c1b, c2b) of class c, pass phi
  int q;                         // parity of the single layer
  int pad;
  int p12[2][kKronMaxTuple];     // c1 | c2 << 8
  int po[2][kKronMaxTuple];      // o1 | o2 << 8
  int j2[2][2][kKronMaxTuple];   // c1b * d_c2 + c2b
};

// Single legs of one site's layer and the table layout of the site.
struct KronSite {
  int d[4] = {1, 1, 1, 1}, e[4] = {1, 1, 1, 1};  // W, N, E, S
  int dphi = 1, ephi = 1;                         // inner index summed between ket and bra
  int no1 = 0, no2 = 0;                           // register tile of the kernel instantiation
  int j1cap = 0, j2cap = 0;
  size_t stage_bytes = 0;                         // one stage's tables; stage k at k * stage_bytes
};

// Chooses the instantiation and table capacities of a layer with these single
// legs; false when no instantiation fits (the dense route is used then).
bool kron_site_layout(const int d[4], const int e[4], int dphi, int ephi, KronSite& ks);
// Tables of every stage for the single layer S (W, N, PHI, E, S; row-major,
// FP64, host) of parity q, into out (kKronStages * ks.stage_bytes bytes).
void kron_build_tables(const KronSite& ks, const double* S, int q, std::vector<char>& out);
// Parity of a single layer: 0 or 1 when every nonzero entry has that parity,
// -1 when both occur (the factorised route needs one).
int kron_layer_parity(const int d[4], const int e[4], int dphi, int ephi, const double* S);

// Rows of one launch: r = slab * nf + f, slab = l * rc + R, with the row
// parity par(l) + par(R) + par(f) (bond splits per sample or fixed).
struct KronRows {
  long long nrows = 0;
  int nf = 1, ef = 1, rc = 1;
  const int* fsplit = nullptr;  // even count of f per sample (else ef)
  const int* lsplit = nullptr;
  int leven = 1;
  const int* rsplit = nullptr;
  int reven = 1;
  // element (c1, slot of c2) of row r of sample s:
  // x + s * xs + slab * sx_slab + f * sxf + c1 * sx1 + slot * sx2
  long long xs = 0, sx_slab = 0, sxf = 0, sx1 = 0, sx2 = 1;
  // output row r of sample s at y + s * ys + r * rowY, element o1 * fold(o2) + slot
  long long ys = 0;
};

// One stage (kKronStages) over the rows. tabs[s] points at the (site, label)
// tables of sample s; with tabs null every sample uses tab_shared.
// rows_per_block <= 0: the default. mfma: the matrix-core kernel where one of
// its instantiations covers the layer (the vector kernel otherwise); groups:
// row groups per block of the matrix-core kernel.
hipError_t launch_kron_rows(const KronSite& ks, int stage, const float* x, float* y, const KronRows& rows,
                            void* const* tabs, int nb, int rows_per_block, hipStream_t st, int groups = 1,
                            const void* tab_shared = nullptr, bool mfma = false);
// Multiply-adds the vector kernel executes per row of a stage (register tile
// included), averaged over the two row parities; kron_mfma_row_macs: the same
// for the matrix-core kernel (its tile padding included).
double kron_row_macs(const KronSite& ks, int stage);
double kron_mfma_row_macs(const KronSite& ks, int stage);
// Raise the LDS limit of every instantiation (once per process, before any
// capture).
void kron_prepare();

}  // namespace spp
