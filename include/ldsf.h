/*
 * ldsf.h -- C-ABI of the MI355X (gfx950) Behler symmetry-function path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (LDSF_OK == 0) and never throws; ldsf_last_error() returns the calling
 * thread's last message.  Host-pointer entry points keep the reference's
 * argument lists and its "+=" semantics into caller-owned buffers
 * (src/SymmetryFunctions.cpp:192,253; python_bindings/python_bindings.cpp:99,
 * 124,147,176,205,241,276 zero them first).  Device-pointer entry points take
 * a plan built once per neighbour list and a hipStream_t passed as void*.
 *
 * Layouts (identical to the reference):
 *   species[n_total]            int32, species code of every row (padding too)
 *   number_of_neighs[n_atoms]   int32, neighbour count of each centre
 *   neighbor_list[sum]          int32, global rows, centres concatenated
 *   coordinates[n_total][3]     float64, row-major
 *   desc[n_atoms][width]        float64, row-major, +=
 *   dE_dzeta[n_atoms][width]    float64
 *   d_coordinates[n_total][3]   float64, +=, padding rows included
 */
#ifndef LDSF_H_
#define LDSF_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDSF_API __attribute__((visibility("default")))

/* status codes */
#define LDSF_OK 0
#define LDSF_ERR_ARG 1             /* bad argument (C++ layer: std::invalid_argument) */
#define LDSF_ERR_DEVICE 2          /* HIP runtime error / no device */
#define LDSF_ERR_UNSUPPORTED 3     /* parameter set or size not supported */
#define LDSF_ERR_FILE 4            /* parameter file I/O or parse (std::runtime_error) */
#define LDSF_ERR_NEIGHBOR 5        /* neighbour-list build failure (collision, huge box) */
#define LDSF_ERR_SINGULAR_CELL 6   /* padding: cell matrix not invertible */
#define LDSF_ERR_NOMEM 7           /* device allocation failed */
#define LDSF_ERR_KIND 8            /* descriptor kind other than SymmetryFunctions */
#define LDSF_ERR_TIMEOUT 9         /* ldsf_server_wait: request not finished in time */

/* precision forms */
#define LDSF_FP64 0       /* plain form: FP64 throughout */
#define LDSF_FP32_BULK 1  /* FP32 exp/cos/pow terms; FP64 geometry and accumulation */

typedef struct ldsf_sf ldsf_sf;          /* a SymmetryFunctions parameter set */
typedef struct ldsf_plan ldsf_plan;      /* device-resident neighbour-list plan */
typedef struct ldsf_neighlist ldsf_neighlist;
typedef struct ldsf_padding ldsf_padding;

LDSF_API const char *ldsf_last_error(void);

/* ------------------------------------------------------------------------
 * Parameter sets
 * ------------------------------------------------------------------------ */

/* SymmetryFunctions(species, cutoff_function, cutoff_matrix, types, sizes,
 * params) (src/SymmetryFunctions.cpp:429-450), reached from
 * DescriptorKind::initDescriptor's SF overload (src/Descriptors.cpp:579-590).
 * cutoff_matrix is n_species x n_species row-major; type_names are "g1".."g5";
 * sizes is [rows0, cols0, rows1, cols1, ...]; params is flat row-major.
 * species_resolved = 1 selects the per-neighbour-species G1-G3 blocks and
 * per-unordered-species-pair G4/G5 blocks (build extension; 0 = reference). */
LDSF_API int ldsf_sf_create(ldsf_sf **out, int n_species, const char *const *species_names,
                            const double *cutoff_matrix, int n_types,
                            const char *const *type_names, const int *sizes,
                            const double *params, int species_resolved);

/* SymmetryFunctions(std::string&) -> initFromFile (src/SymmetryFunctions.cpp:262-407).
 * Species are sorted as the reference's std::set sorts them. */
LDSF_API int ldsf_sf_create_from_file(ldsf_sf **out, const char *path);

LDSF_API void ldsf_sf_destroy(ldsf_sf *sf);

/* DescriptorKind::width (include/Descriptors.hpp:184) */
LDSF_API int ldsf_sf_width(const ldsf_sf *sf);
/* number of parameters: sum of rows x cols over the blocks (the length of
 * initDescriptor's flat params vector) */
LDSF_API int ldsf_sf_num_params(const ldsf_sf *sf);
LDSF_API int ldsf_sf_num_species(const ldsf_sf *sf);
/* species name k (sorted order for file init); NULL if out of range */
LDSF_API const char *ldsf_sf_species_name(const ldsf_sf *sf, int k);
/* default precision of the reference-signature entry points below */
LDSF_API int ldsf_sf_set_precision(ldsf_sf *sf, int precision);

/* ------------------------------------------------------------------------
 * Reference-signature entry points (host pointers)
 * ------------------------------------------------------------------------ */

/* Descriptor::compute (include/Descriptors.hpp:130-132, src/Descriptors.cpp:121-136) */
LDSF_API int ldsf_compute(ldsf_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                          const int *number_of_neighs, const double *coordinates, double *desc);

/* Descriptor::compute_single_atom (Descriptors.hpp:163-165, Descriptors.cpp:355-365) */
LDSF_API int ldsf_compute_single_atom(ldsf_sf *sf, int index, int n_atoms, const int *species,
                                      const int *neighbor_list, int number_of_neighs,
                                      const double *coordinates, double *desc);

/* Descriptor::compute_batch (Descriptors.hpp:147-149, Descriptors.cpp:138-163) */
LDSF_API int ldsf_compute_batch(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                const int *config_ptr, const int *species,
                                const int *neighbor_list, const int *number_of_neighs,
                                const double *coordinates, double *desc);

/* ldsf_compute / ldsf_compute_batch with desc overwritten (=) instead of
 * accumulated (+=), so desc's old contents are not uploaded: for callers whose
 * desc starts at zero, such as the Python module, which returns a new zeroed
 * array (the reference's binding zero-fills it, python_bindings.cpp:146-147).
 * No reference counterpart. */
LDSF_API int ldsf_compute_overwrite(ldsf_sf *sf, int n_atoms, const int *species,
                                    const int *neighbor_list, const int *number_of_neighs,
                                    const double *coordinates, double *desc);
LDSF_API int ldsf_compute_batch_overwrite(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                          const int *config_ptr, const int *species,
                                          const int *neighbor_list, const int *number_of_neighs,
                                          const double *coordinates, double *desc);

/* Descriptor::gradient (Descriptors.hpp:58-60, Descriptors.cpp:165-190).
 * d_coordinates += J^T dE_dzeta, and desc += the descriptors: the reference's
 * Enzyme call runs the forward too (enzyme_dup desc), so the same holds here,
 * from one fused pass; desc may be NULL (the reference's callers pass a
 * throwaway copy, python_bindings.cpp:236-238).  The same holds for the
 * single-atom and batch forms below. */
LDSF_API int ldsf_gradient(ldsf_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                           const int *number_of_neighs, const double *coordinates,
                           double *d_coordinates, double *desc, const double *dE_dzeta);

/* Descriptor::gradient_single_atom (Descriptors.hpp:97-99, Descriptors.cpp:368-395) */
LDSF_API int ldsf_gradient_single_atom(ldsf_sf *sf, int index, int n_atoms, const int *species,
                                       const int *neighbor_list, int number_of_neighs,
                                       const double *coordinates, double *d_coordinates,
                                       double *desc, const double *dE_dzeta);

/* Descriptor::gradient_batch (Descriptors.hpp:78-80, Descriptors.cpp:263-292) */
LDSF_API int ldsf_gradient_batch(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                 const int *config_ptr, const int *species,
                                 const int *neighbor_list, const int *number_of_neighs,
                                 const double *coordinates, double *d_coordinates, double *desc,
                                 const double *dE_dzeta);

/* Descriptor::jacobian (Descriptors.hpp:117-118, Descriptors.cpp:454-486):
 * J_coordinates[width*n_atoms][3*n_total_atoms] += d desc / d coords. */
LDSF_API int ldsf_jacobian(ldsf_sf *sf, int n_atoms, int n_total_atoms, const int *species,
                           const int *neighbor_list, const int *number_of_neighs,
                           const double *coordinates, double *J_coordinates);

/* Hyperparameter gradients: what the reference's gradient computes into the
 * Enzyme shadow of the descriptor object and then deletes unread
 * (src/Descriptors.cpp:176-188, enzyme_dup desc_kind, d_desc_kind).
 * dparams[ldsf_sf_num_params] += dE/d(parameters) in initDescriptor's flat
 * layout (per block rows x cols, row-major: g2 (eta, Rs), g3 (kappa), g4/g5
 * (zeta, lambda, eta); g1 has none); drcut[n_species^2] += dE/d(cutoff matrix)
 * (rcut_2D_, row-major; NULL: not returned).  FP64; at most 8 species. */
LDSF_API int ldsf_param_gradient(ldsf_sf *sf, int n_atoms, const int *species,
                                 const int *neighbor_list, const int *number_of_neighs,
                                 const double *coordinates, const double *dE_dzeta,
                                 double *dparams, double *drcut);

/* Descriptor::num_gradient_single_atom (Descriptors.hpp:101-103,
 * Descriptors.cpp:488-518, include/finite_difference.hpp:46-143):
 * d_coordinates[0:3] += sum_w dE_dzeta[w] d desc_index[w] / d x_index, by the
 * order-8 central stencil with h = (551.25 eps)^(1/9); the descriptor
 * evaluations run on the GPU.  n_atoms is the number of coordinate rows. */
LDSF_API int ldsf_num_gradient_single_atom(ldsf_sf *sf, int index, int n_atoms,
                                           const int *species, const int *neighbor_list,
                                           int number_of_neighs, const double *coordinates,
                                           double *d_coordinates, const double *dE_dzeta);

/* ------------------------------------------------------------------------
 * Device-pointer entry points (no reference counterpart: the reference has
 * no device memory).  A plan holds the CSR offsets of the neighbour list, its
 * transpose (for the atomic-free adjoint gather) and scratch, on the current
 * HIP device.  All device pointers must be on that device; work is ordered on
 * `stream` (a hipStream_t, NULL = default stream).
 * ------------------------------------------------------------------------ */

/* Plan over centres centers[0..n_centers) (device int32, copied; NULL = rows
 * 0..n_centers-1), with n_total coordinate rows.  Synchronises `stream` once
 * (to size the launch from the largest neighbour count).  The plan keeps sf,
 * d_species and d_neighbor_list without copying them: they must outlive it. */
LDSF_API int ldsf_plan_create(ldsf_plan **out, const ldsf_sf *sf, int n_centers,
                              const int *d_centers, int n_total, const int *d_species,
                              const int *d_neighbor_list, const int *d_number_of_neighs,
                              void *stream);

/* Plan over the concatenated configurations of compute_batch/gradient_batch:
 * n_atoms[] and config_ptr[] are HOST arrays (Descriptors.cpp:154-161). */
LDSF_API int ldsf_plan_create_batch(ldsf_plan **out, const ldsf_sf *sf, int n_configurations,
                                    const int *n_atoms, const int *config_ptr, int n_total,
                                    const int *d_species, const int *d_neighbor_list,
                                    const int *d_number_of_neighs, void *stream);

LDSF_API void ldsf_plan_destroy(ldsf_plan *plan);
LDSF_API int ldsf_plan_num_centers(const ldsf_plan *plan);
LDSF_API long ldsf_plan_num_edges(const ldsf_plan *plan);
LDSF_API int ldsf_plan_max_neighbors(const ldsf_plan *plan);

/* d_desc[n_centers][width] += descriptors of the plan's centres. */
LDSF_API int ldsf_plan_compute(ldsf_plan *plan, const double *d_coordinates, double *d_desc,
                               int precision, void *stream);

/* d_dcoords[n_total][3] += J^T d_dE_dzeta (dE_dzeta is [n_centers][width]). */
LDSF_API int ldsf_plan_gradient(ldsf_plan *plan, const double *d_coordinates,
                                const double *d_dE_dzeta, double *d_dcoords, int precision,
                                void *stream);

/* The reference's gradient semantics on the device: d_desc[n_centers][width]
 * += descriptors and d_dcoords[n_total][3] += J^T d_dE_dzeta, from one fused
 * kernel that shares the pair and triplet work between the two (then the
 * gather).  ldsf_plan_forward_adjoint_edges is its first half, the fused
 * kernel alone; ldsf_plan_gather completes it. */
LDSF_API int ldsf_plan_forward_gradient(ldsf_plan *plan, const double *d_coordinates,
                                        const double *d_dE_dzeta, double *d_desc,
                                        double *d_dcoords, int precision, void *stream);
LDSF_API int ldsf_plan_forward_adjoint_edges(ldsf_plan *plan, const double *d_coordinates,
                                             const double *d_dE_dzeta, double *d_desc,
                                             int precision, void *stream);
/* No reference counterpart: ldsf_plan_forward_gradient with the outputs
 * assigned instead of added, d_desc = descriptors and d_dcoords = J^T
 * d_dE_dzeta (every entry written; bitwise the += into zeroed arrays), for
 * callers that want fresh outputs without clearing them first. */
LDSF_API int ldsf_plan_forward_gradient_assign(ldsf_plan *plan, const double *d_coordinates,
                                               const double *d_dE_dzeta, double *d_desc,
                                               double *d_dcoords, int precision, void *stream);
/* Its two halves (no reference counterpart): the fused kernel with d_desc
 * assigned, and the gather with d_dcoords assigned (rows without records
 * set to 0). */
LDSF_API int ldsf_plan_forward_adjoint_edges_assign(ldsf_plan *plan, const double *d_coordinates,
                                                    const double *d_dE_dzeta, double *d_desc,
                                                    int precision, void *stream);
LDSF_API int ldsf_plan_gather_assign(ldsf_plan *plan, double *d_dcoords, void *stream);

/* ldsf_param_gradient on the device: d_dparams[ldsf_sf_num_params] and
 * d_drcut[n_species^2] (may be NULL) +=.  Allocates its partial sums and
 * synchronises `stream`. */
LDSF_API int ldsf_plan_param_gradient(ldsf_plan *plan, const double *d_coordinates,
                                      const double *d_dE_dzeta, double *d_dparams,
                                      double *d_drcut, void *stream);

/* The two halves of ldsf_plan_gradient.  ldsf_plan_adjoint_edges runs the
 * adjoint kernel and keeps one dE/dr vector per neighbour-list edge and per
 * centre in the plan's workspace; ldsf_plan_gather then adds, for every
 * coordinate row, its vectors (in ascending edge order, so the sum is the same
 * on every run) into d_dcoords[n_total][3].  Calls on one plan must not
 * overlap. */
LDSF_API int ldsf_plan_adjoint_edges(ldsf_plan *plan, const double *d_coordinates,
                                     const double *d_dE_dzeta, int precision, void *stream);
LDSF_API int ldsf_plan_gather(ldsf_plan *plan, double *d_dcoords, void *stream);

/* The adjoint's record buffer (no reference counterpart).  A plan allocates
 * 3 * ldsf_plan_num_records(plan) doubles on its first adjoint call and frees
 * them with the plan; ldsf_plan_set_records hands it a caller's device buffer
 * of at least that size instead (kept, not freed, by the plan; it must
 * outlive it), so that callers who build a plan per call can take the records
 * from their own memory pool.  d_records must be device memory on the plan's
 * device, float64 (n_doubles counts doubles) and contiguous; only the count
 * can be checked here. */
LDSF_API long ldsf_plan_num_records(const ldsf_plan *plan);
LDSF_API int ldsf_plan_set_records(ldsf_plan *plan, double *d_records, long n_doubles);

/* Which kernel family a plan uses: 0 = general (any parameter set), 1 = wave-
 * per-centre LDS kernels.  Setting 0 forces the general kernels. */
LDSF_API int ldsf_plan_kernel_family(const ldsf_plan *plan);
LDSF_API int ldsf_plan_set_kernel_family(ldsf_plan *plan, int family);

/* ------------------------------------------------------------------------
 * Allocation-free gradient (no reference counterpart).  For callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  ldsf_gradient's semantics
 * (Descriptor::gradient: d_d_coordinates[n_total][3] += J^T d_dE_dzeta and,
 * when d_desc is not NULL, d_desc[n_atoms][width] += the descriptors) on
 * device pointers in the reference's layouts, centres = rows 0..n_atoms-1.
 * It calls no hipMalloc, hipFree or allocating library routine: the
 * parameter tables, CSR offsets, the transposed neighbour list, the edge
 * records and flags live in one static device array of
 * LDSF_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held until the process
 * ends.  A call needs about 32 (n_neigh + n_atoms) + 12 n_total + 8 n_atoms
 * bytes of it plus the parameter tables (64 MiB: about 1.9 M neighbour-list
 * entries); a larger input is refused with LDSF_ERR_UNSUPPORTED before any
 * device work.  n_neigh = sum(number_of_neighs) and max_neighbors bounds every
 * number_of_neighs[i]; both are checked on the device, with the species codes
 * and neighbour indices, before any kernel reads through them: a rejected
 * call (LDSF_ERR_ARG) leaves d_desc and d_d_coordinates unchanged.  Work is
 * ordered on `stream`; the call synchronises it once at the end, to return
 * the device-side checks.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU race. */
#define LDSF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
LDSF_API int ldsf_gradient_noalloc(ldsf_sf *sf, int n_atoms, int n_total, int n_neigh,
                                   int max_neighbors, const int *d_species,
                                   const int *d_neighbor_list, const int *d_number_of_neighs,
                                   const double *d_coordinates, double *d_d_coordinates,
                                   double *d_desc, const double *d_dE_dzeta, int precision,
                                   void *stream);

/* ------------------------------------------------------------------------
 * Serving loop (no reference counterpart).  Requests of descriptors + dE/dr
 * over device-resident inputs are admitted in arrival order into lanes --
 * device buffers for one batch each, sized by the config, as many lanes as
 * hbm_budget_bytes holds (at most max_lanes) -- and each batch runs as one
 * replay of the lane's hipGraph on the lane's stream, concurrently with the
 * other lanes.  Requests arriving meanwhile wait only for a free lane.
 * Outputs are overwritten (=), not accumulated.  A request's buffers must stay
 * valid and unmodified until ldsf_server_wait has returned for it.
 * ------------------------------------------------------------------------ */
typedef struct ldsf_server ldsf_server;

typedef struct {
    int precision;            /* LDSF_FP64 or LDSF_FP32_BULK */
    size_t hbm_budget_bytes;  /* device memory the lanes may hold together */
    int lane_centers;         /* centres per batch */
    int lane_rows;            /* coordinate rows (padding included) per batch */
    int lane_edges;           /* neighbour-list entries per batch */
    int max_requests;         /* requests per batch (<= 65535) */
    int max_lanes;            /* upper bound on concurrent batches */
    int dispatch_min_centers; /* > 0: a free lane starts once this many centres are queued */
    int batch_window_us;      /* > 0: ... or once the oldest queued request has waited this
                                 long (microseconds).  With no window, a free lane starts at
                                 once (no threshold) or when the device is idle */
} ldsf_server_config;

typedef struct {
    int n_atoms;                   /* centres: rows 0..n_atoms-1 */
    int n_total;                   /* coordinate rows */
    int n_neigh;                   /* entries of neighbor_list = sum(number_of_neighs) */
    int max_neighbors;             /* bound on every number_of_neighs[i] */
    const int *species;            /* device [n_total] */
    const int *number_of_neighs;   /* device [n_atoms] */
    const int *neighbor_list;      /* device [n_neigh], rows 0..n_total-1 */
    const double *coordinates;     /* device [n_total][3] */
    const double *dE_dzeta;        /* device [n_atoms][width]; NULL = zero (no gradient) */
    double *desc;                  /* device [n_atoms][width], written */
    double *d_coordinates;         /* device [n_total][3], written; NULL = not returned */
} ldsf_request;

/* Create on the current HIP device; starts one worker and one completion
 * thread per lane.  The server uses sf's device tables: sf must outlive it. */
LDSF_API int ldsf_server_create(ldsf_server **out, const ldsf_sf *sf,
                                const ldsf_server_config *config);
LDSF_API int ldsf_server_num_lanes(const ldsf_server *server);
/* Queue a request (thread-safe); *ticket identifies it for ldsf_server_wait. */
LDSF_API int ldsf_server_submit(ldsf_server *server, const ldsf_request *request,
                                long long *ticket);
/* Block until the request has finished (timeout_s < 0: no limit); returns its
 * status (LDSF_ERR_NEIGHBOR: an index, species code or count out of range was
 * found on the device) and the time from submission to completion.  A ticket
 * is waited for once. */
LDSF_API int ldsf_server_wait(ldsf_server *server, long long ticket, double timeout_s,
                              double *latency_s);
LDSF_API int ldsf_server_stats(const ldsf_server *server, long long *batches,
                               long long *requests, long long *atoms, long long *graphs);
/* Finishes every queued request, then frees the lanes. */
LDSF_API void ldsf_server_destroy(ldsf_server *server);

/* ------------------------------------------------------------------------
 * Input-layout producers (python_bindings/neighbor/neighbor_list.h)
 * ------------------------------------------------------------------------ */

/* nbl_create_paddings (neighbor_list.cpp:283-418) */
LDSF_API int ldsf_nbl_create_paddings(int n_particles, double cutoff, const double *cell,
                                      const int *pbc, const double *coords, const int *species,
                                      ldsf_padding **out);
LDSF_API int ldsf_padding_count(const ldsf_padding *p);
LDSF_API int ldsf_padding_copy(const ldsf_padding *p, double *coords, int *species, int *image);
LDSF_API void ldsf_padding_free(ldsf_padding *p);

/* nbl_initialize / nbl_build / nbl_clean (neighbor_list.cpp:61-243) */
LDSF_API ldsf_neighlist *ldsf_nbl_create(void);
LDSF_API void ldsf_nbl_free(ldsf_neighlist *nl);
LDSF_API int ldsf_nbl_build(ldsf_neighlist *nl, int n_particles, const double *coords,
                            double influence_distance, int n_cutoffs, const double *cutoffs,
                            const int *need_neighbors);
/* nbl_get_neigh (neighbor_list.cpp:246-280), KIM GetNeighbor signature; returns 1 on error */
LDSF_API int ldsf_nbl_get_neigh(void const *data_object, int n_cutoffs, double const *cutoffs,
                                int neighbor_list_index, int particle_number,
                                int *number_of_neighbors, int const **neighbors_of_particle);
LDSF_API int ldsf_nbl_num_lists(const ldsf_neighlist *nl);
LDSF_API double ldsf_nbl_list_cutoff(const ldsf_neighlist *nl, int k);
LDSF_API int ldsf_nbl_num_particles(const ldsf_neighlist *nl);
LDSF_API long ldsf_nbl_total(const ldsf_neighlist *nl, int k, int n_query);
/* NeighborList.get_numneigh_and_neighlist_1D (python_package/libdescriptor/neighbor.py:108-154) */
LDSF_API int ldsf_nbl_flatten(const ldsf_neighlist *nl, int k, int n_query, int *numneigh,
                              int *neighlist);

/* ------------------------------------------------------------------------
 * Device-resident layout producer (GPU twin of the producers above)
 * ------------------------------------------------------------------------ */
typedef struct ldsf_layout ldsf_layout;

/* NeighborList(conf, infl_dist) (python_package/libdescriptor/neighbor.py:35-78)
 * built on the current device: padding atoms as nbl_create_paddings
 * (neighbor_list.cpp:283-418) appended after the n_contrib contributing rows,
 * then the full neighbour list of the contributing rows as nbl_build
 * (neighbor_list.cpp:78-243, needNeighbors = 1 for contributing rows only,
 * one cutoff = infl_dist) in the flat layout of get_numneigh_and_neighlist_1D
 * (neighbor.py:108-154).  Same rows, order and FP64 values as the host
 * producers.  cell[9] (rows = cell vectors) and pbc[3] are host arrays;
 * d_coords[n_contrib][3] (float64) and d_species[n_contrib] (int32) are device
 * arrays.  Errors as the host producers: LDSF_ERR_SINGULAR_CELL,
 * LDSF_ERR_NEIGHBOR (collision, too many cells).  Synchronises `stream`. */
LDSF_API int ldsf_layout_create(ldsf_layout **out, int n_contrib, double cutoff,
                                const double *cell, const int *pbc, const double *d_coords,
                                const int *d_species, void *stream);
LDSF_API void ldsf_layout_destroy(ldsf_layout *layout);
/* sizes: rows = n_contrib + n_padding; n_neigh = sum(numneigh) */
LDSF_API int ldsf_layout_info(const ldsf_layout *layout, int *n_contrib, int *n_padding,
                              long *n_neigh, int *max_neighbors);
/* device arrays owned by the layout: coords[rows][3], species[rows],
 * image[rows] (NeighborList.get_image: a contributing row's image is itself),
 * numneigh[n_contrib], neighlist[n_neigh] */
LDSF_API int ldsf_layout_arrays(const ldsf_layout *layout, const double **coords,
                                const int **species, const int **image, const int **numneigh,
                                const int **neighlist);
/* copies the layout's arrays into caller-owned device buffers (any may be NULL)
 * of the sizes ldsf_layout_info reports, ordered on `stream` */
LDSF_API int ldsf_layout_copy(const ldsf_layout *layout, double *d_coords, int *d_species,
                              int *d_image, int *d_numneigh, int *d_neighlist, void *stream);
/* assemble_forces(forces, n, padding_image) (neighbor.py:227-263) on the
 * device: d_out[n_contrib][3] = the contributing rows of d_forces[rows][3]
 * plus the rows of their padding images, summed in the reference's order
 * (per atom 0 + sum of its image rows, then added, when n < n_padding; else
 * the image rows added one by one in padding order). */
LDSF_API int ldsf_layout_assemble_forces(const ldsf_layout *layout, const double *d_forces,
                                         double *d_out, void *stream);
/* ldsf_plan_gather into a zeroed d_coordinates followed by
 * ldsf_layout_assemble_forces, in one pass and bitwise the same:
 * d_out[n_contrib][3] = each contributing atom's dE/dr with its padding rows
 * folded on in the reference's order (neighbor.py:227-263), from the records
 * of the plan's last adjoint call.  The plan must be over the layout's rows
 * (ldsf_layout_arrays); the padding rows' dE/dr is never written. */
LDSF_API int ldsf_plan_gather_fold(ldsf_plan *plan, const ldsf_layout *layout, double *d_out,
                                   void *stream);
/* the same fold for any padding_image (device int32 [n_padding], values in
 * [0, n_contrib)); d_forces[n_contrib + n_padding][3].  Synchronises `stream`. */
LDSF_API int ldsf_assemble_forces(int n_contrib, int n_padding, const double *d_forces,
                                  const int *d_padding_image, double *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LDSF_H_ */
