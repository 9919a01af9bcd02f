// Launchers of the HIP kernels (host-callable).
#ifndef LDSF_SF_KERNELS_H_
#define LDSF_SF_KERNELS_H_

#include <hip/hip_runtime.h>

#include "ldsf_internal.h"

struct ldsf_layout;

namespace ldsf {

// Sizes of a launch read on the device, for launches replayed from a graph
// with different batches (the serving loop): hdr[0] centres, hdr[1]
// coordinate rows, hdr[2] neighbour-list edges.  With req_flag set, centres
// whose request (req_of[c]) is flagged are skipped: they contribute nothing
// and their edge records are left unwritten.  A default DevCounts (all null)
// means the sizes passed as arguments.
struct DevCounts {
    const int *hdr = nullptr;
    const int *req_of = nullptr;
    const int *req_flag = nullptr;
    int max_blocks = 0;  // wave kernels: cap on the persistent grid (0 = 16384)
    // per centre: its descriptor row, written with = instead of += (null: desc + c * width)
    double *const *out_row = nullptr;
    // per centre: its dE/dzeta row (null: dE + c * width)
    const double *const *in_row = nullptr;
};

// sf_general.hip: one thread per centre, any parameter set.  With dc.hdr set,
// n_centers (and n_edges) give the grid's capacity; the launch covers hdr[0].
hipError_t launch_fwd_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              double *desc, hipStream_t s, DevCounts dc = DevCounts());
hipError_t launch_adj_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              const double *dE, const int *inv_pos, double *edge, int n_edges,
                              hipStream_t s, DevCounts dc = DevCounts());

// Hyperparameter gradients (FP64, general kernels): dparams[n_params] +=
// dE/d(parameters) in initDescriptor's flat layout, drcut[ns*ns] += dE/d(rcut
// matrix) (NULL: not returned), for the centres' dE/dzeta rows.  Allocates its
// partials and synchronises `s`.  At most 8 species.
hipError_t launch_param_gradient(const SFTab &tab, int n_centers, const int *centers,
                                 const int *offs, const int *nl, const int *species,
                                 const double *x, const double *dE, double *dparams,
                                 double *drcut, hipStream_t s);

// sf_wave.hip: one wavefront per centre, neighbours staged in LDS.  They
// cover parameter sets whose two-body rows are g2 (at most 8) and whose
// three-body rows are g4 forming a grid zeta x lambda x eta (at most 4 x 2 x 2
// distinct values), in the reference layout or the two-species resolved one.
// Row parameters of the wave kernels; one copy in device memory per plan,
// staged into LDS by every workgroup (uniform values read from LDS instead of
// occupying scalar registers across the loops).
struct WaveParams {
    double eta2[8], rs2[8];      // g2 rows
    double zeta[4], coef[4];     // distinct g4 zeta and 2^(1 - zeta)
    double lam[2], eta3[2];      // distinct g4 lambda and eta
    double lp[2], lm[2];         // (1 + lambda) / 2 and (1 - lambda) / 2
    int izeta[4];                // integer zeta or -1
    int lamcode[2];              // +1 / -1 for lambda = +1 / -1, else 0
    int nsq;                     // squarings needed for the largest integer zeta
    int slots[64];               // [0,16): g2 slot of k = s*8+q; [16,64): g4 slot of
                                 // p*16 + (z*2+l)*2+e; -1 unused
    float e2l[8], m2e[8], rs2f[8];  // FP32-bulk g2 rows: eta log2(e), -2 eta, Rs (0 past
                                    // the set's rows)
};
struct WaveTab {
    int n_species;
    int nr2;                // g2 rows (<= 8)
    int nz, nl, ne;         // distinct zeta, lambda, eta of the g4 grid
    int nsb, npb;           // species blocks (1 or 2) and pair blocks (1 or 3)
    int has3;               // any g4 row
    int rc_uniform;         // every species pair has the same cutoff
    int zgrid;              // zeta grid (1, 2, 4, 16) = 1, (1, 2, 4, 8) = 2, else 0: the FP32-bulk
                            // kernels' compile-time squaring chains
    int fp32_exp_ok;        // every exponent the FP32-bulk form takes in FP32 is <= kFp32ExpMax
    int nl_species;         // the neighbour-list entries carry the neighbour's species code
                            // (entry = row | species << kNlSpeciesShift, a plan's packed list)
    const double *rc;       // device, n_species^2
    const WaveParams *par;  // device
};
constexpr int kWaveSlots = 64;
constexpr int kMaxWaveSpecies = 8;
// A plan's packed neighbour list for the wave kernels: the row in the low 28
// bits, the species code (< kMaxWaveSpecies) above, so entries stay positive.
constexpr int kNlSpeciesShift = 28;
constexpr int kNlIndexMask = (1 << kNlSpeciesShift) - 1;
// Largest exponent eta (r - Rs)^2 (g2) or eta (r_ij^2 + r_ik^2 + r_jk^2) (g4)
// that the FP32-bulk wave kernels form in FP32; sets that reach beyond it
// over the cutoff sphere run the FP32-bulk general kernels, which split the
// exponent in FP64 (relative accuracy ~1e-7 for any exponent).
constexpr double kFp32ExpMax = 32.0;
// whether the wave kernels (when built for the set) run in this precision form
inline bool wave_runs(const WaveTab &wt, int precision)
{
    return precision == 0 || wt.fp32_exp_ok != 0;
}
// host: true (and the table and parameters) if the wave kernels cover this
// parameter set
bool wave_build(const ldsf_sf *sf, WaveTab *wt, WaveParams *par);
// Lanes per centre (16, 32 or 64) for the largest neighbour count and number
// of species; 0 if the wave kernels cannot run (more than 64 neighbours, or
// too much LDS).
int wave_group_size(int max_nn, int n_species);
size_t wave_lds_bytes(int max_nn, bool adjoint, int n_species, int group);
constexpr size_t kWaveMaxLds = 64 * 1024;
hipError_t launch_fwd_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           double *desc, int width, int max_nn, int group, hipStream_t s,
                           DevCounts dc = DevCounts());
// The adjoint kernels write the vector of edge e (and of centre c, as edge
// n_edges + c) at record inv_pos[e] of `edge`, the order the gather reads.
// With desc (or dc.out_row) set, the fused kernel runs: it also adds each
// centre's descriptors into desc (or assigns them to dc.out_row[c]), from the
// same pair and triplet work.
// assign (fused only): desc[c] = the descriptors instead of +=, as dc.out_row
// does per centre.
hipError_t launch_adj_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           const double *dE, const int *inv_pos, double *edge, int n_edges,
                           int width, int max_nn, int group, hipStream_t s,
                           DevCounts dc = DevCounts(), double *desc = nullptr,
                           bool assign = false);
// the FP64 form's launches, in units of their own for their compiler options
// (sf_wave_fp64.hip: forward and adjoint; sf_wave_fp64_fused.hip: the fused
// kernel); launch_fwd_wave and launch_adj_wave call them for precision 0
hipError_t launch_fwd_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                const int *offs, const int *nl, const int *species, const double *x,
                                double *desc, int width, int max_nn, int group, hipStream_t s,
                                DevCounts dc);
hipError_t launch_adj_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                const int *offs, const int *nl, const int *species, const double *x,
                                const double *dE, const int *inv_pos, double *edge, int n_edges,
                                int width, int max_nn, int group, hipStream_t s, DevCounts dc,
                                double *desc, bool assign);
hipError_t launch_fused_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                  const int *offs, const int *nl, const int *species,
                                  const double *x, const double *dE, const int *inv_pos,
                                  double *edge, int n_edges, int width, int max_nn, int group,
                                  hipStream_t s, DevCounts dc, double *desc, bool assign);

// sf_plan.hip: neighbour-list plan building and the adjoint gather
hipError_t launch_plan_validate(int n_centers, const int *centers, int n_total, const int *numneigh,
                                const int *species, int n_species, int *flags, hipStream_t s);
hipError_t launch_edge_targets_count(long n_edges, int n_centers, const int *nl, const int *centers,
                                     int n_total, int *count, int *flags, hipStream_t s);
// skips (as the count flags) targets outside [0, n_total)
hipError_t launch_edge_targets_fill(long n_edges, int n_centers, const int *nl, const int *centers,
                                    int n_total, const int *in_start, int *cursor, int *in_edges,
                                    hipStream_t s);
// The transposed list in one go, by buckets of 256 rows (no per-entry global
// atomics): in_start[0..n_total], in_edges (ids ascending per row) and
// inv_pos, the same arrays as count + scan + fill + sort.  Needs
// bucket_count_of(n_total) <= 4096 (bucketed_fits); bcount and bbase hold
// nb + 1 ints, bcursor nb, stage n_edges + n_centers pairs; tmp is
// launch_scan_int's scratch for nb entries.  Out-of-range targets set flags
// bit 1 and are left out, as the count does.
int bucket_count_of(int n_total);
bool bucketed_fits(int n_total);
hipError_t launch_transpose_bucketed(long n_edges, int n_centers, const int *nl,
                                     const int *centers, int n_total, int *bcount, int *bbase,
                                     int *bcursor, int2 *stage, int *in_start, int *in_edges,
                                     int *inv_pos, int *flags, void *tmp, size_t tmp_bytes,
                                     hipStream_t s);
// sorts every row's bucket of in_edges and writes inv_pos[edge id] = its position
hipError_t launch_sort_inverse(int n_total, const int *in_start, int *in_edges, int *inv_pos,
                               hipStream_t s);
// dcoords[a] += the sum of row a's records (assign: = it, rows without records
// 0); nothing when *skip != 0 (skip may be NULL)
hipError_t launch_gather(int n_total, const int *in_start, const double *edge, double *dcoords,
                         hipStream_t s, const int *skip = nullptr, bool assign = false);
hipError_t launch_max_int(int n, const int *v, int *out, hipStream_t s);
// out[e] = nl[e] | species[nl[e]] << kNlSpeciesShift for validated entries
hipError_t launch_pack_species(long n_edges, const int *nl, const int *species, int *out,
                               hipStream_t s);
// v[0 .. n) = 0 by a kernel (no memset node where a caller captures a graph)
hipError_t launch_zero_f64(long n, double *v, hipStream_t s);

// sf_layout.hip: a device layout's rows, and the gather of adjoint records
// (row r's records at in_start[r] .. in_start[r+1]) folded onto its
// contributing atoms (d_out[n_contrib][3]), as the gather into a zeroed
// d_coordinates followed by ldsf_layout_assemble_forces, bitwise
int layout_rows(const ldsf_layout *layout);
int layout_gather_fold(const ldsf_layout *layout, const int *in_start, const double *edge,
                       double *d_out, hipStream_t s);
hipError_t launch_scan_int(int n, const int *in, int *out_exclusive_plus_total, void *tmp,
                           size_t *tmp_bytes, hipStream_t s);

}  // namespace ldsf

#endif  // LDSF_SF_KERNELS_H_
