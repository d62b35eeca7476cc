/*
 * oops_cce.h — C ABI of the CCE host pipeline around the expv hot path.
 *
 * These entry points build what the reference's CCE driver builds on the
 * host and feed the batched device expv of oops_expv.h.  Reference
 * interfaces replaced (paths relative to the reference's src/):
 *   oops_nv_states           <- NVCenter::make_espin_hamiltonian + eig_sym
 *                               (source/app/DefectCenter.cpp:32-45)
 *   oops_cce_cluster_terms   <- EnsembleCCE::create_spin_hamiltonian /
 *                               create_spin_liouvillian / create_incoherent_operator /
 *                               create_spin_density_state
 *                               (source/app/ensemble_cce.cpp:57-102), flattened as
 *                               MatExpVector does (source/math/MatExp.cpp:219-262)
 *   oops_clusters_enumerate  <- cSpinCollection::make + getConnectionMatrix +
 *                               cDepthFirstPathTracing::generate
 *                               (source/spin/SpinCollection/SpinCollection.cpp:31-57,
 *                                source/spin/SpinCluster/SpinClusterAlgorithm.cpp:98-148)
 *   oops_clusters_parents    <- cClusterIndex sub-cluster positions
 *                               (SpinClusterAlgorithm.cpp:50-63, SpinClusterIndex.cpp:29-42)
 *   oops_cce_reduce          <- CCE::cce_coherence_reduction + compuate_final_coherence
 *                               (source/app/cce.cpp:151-189)
 *   oops_cce_hahn_batch_create <- CCE::run_each_clusters (source/app/cce.cpp:54-70)
 *                               with EnsembleCCE::cluster_evolution for a Hahn echo
 *                               (source/app/ensemble_cce.cpp:33-55) on the device expv
 *   oops_bath_states, oops_cce_single_sample_* <- SingleSampleCCE
 *                               (source/app/single_sample_cce.cpp:35-128)
 *   oops_bath_diamond        <- no reference counterpart: the seeded synthetic
 *                               13C bath of the benchmark workloads.
 */
#ifndef OOPS_CCE_H
#define OOPS_CCE_H

#include <stdint.h>

#include "oops_expv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oops_terms oops_terms;
typedef struct oops_clusters oops_clusters;

/* Parameters of EnsembleCCE (EnsNVnBath's option set, SI units / Angstrom). */
typedef struct {
    double B[3];                /* magnetic field, T */
    double nv_pos[3];           /* nitrogen position, Angstrom; electron at + 3.57/4 (1,1,1) */
    int32_t state0, state1;     /* NV eigenstate indices (ascending energy) */
    double dephasing_rate;      /* Hz (times 2 pi internally); 0 drops the terms */
    double dephasing_axis[3];
    double pol[3];              /* bath polarisation of the initial state */
} oops_cce_params;

/* Seeded diamond lattice (8-site cubic cell of side a, Angstrom), sites of
 * the NV (origin and a/4 (1,1,1)) excluded, each other site within the
 * region occupied with probability p (splitmix64 stream of `seed`, one draw
 * per candidate site in cell-major order).  shape 0: sphere |r| <= radius,
 * 1: cube max|x_i| <= radius.  Writes up to max_n coordinates (x,y,z). */
int oops_bath_diamond(double a, double radius, double p, uint64_t seed, int32_t shape, double* coords,
                      int32_t max_n, int32_t* n_out);

/* Spin-coordinate files (cSpinSourceFromFile::generate and
 * cSpinSource::export_coordinates, SpinSource.cpp:20-79): first the spin
 * count, then one line per spin "isotope x y z" (Angstrom).  Reading writes
 * up to max_n spins (coords x,y,z; isotope names NUL-terminated in 16-byte
 * slots, may be NULL) and the count in *n_out (OOPS_ERR_CAPACITY when the file
 * holds more).  Writing uses the reference's format: tab-separated, default
 * six significant digits. */
int oops_read_xyz(const char* path, int32_t max_n, double* coords, char* isotopes, int32_t* n_out);
int oops_write_xyz(const char* path, int32_t n, const double* coords, const char* isotopes);

/* NV eigenvalues (rad/s, ascending) and eigenvectors (3x3 complex,
 * column-major, interleaved). */
int oops_nv_states(const double* pos, const double* B, double* evals, double* evecs);

/* Term table of one cluster of k 13C spins: which = 0 lvA, 1 lvB,
 * 2 -lvB^dagger, 3 H0, 4 H1, 5 -lvA^dagger.  rho (n complex) and D may be NULL. */
int oops_cce_cluster_terms(const double* coords, int32_t k, const oops_cce_params* P, int32_t which,
                           oops_terms** out, double* rho, double* D);
int oops_terms_view(const oops_terms* t, oops_term_table* view);
int oops_terms_sizes(const oops_terms* t, int64_t* nterm, int64_t* npos, int64_t* nmat);
int oops_terms_destroy(oops_terms* t);

/* Connected clusters of 1..max_order spins for the pair graph d <= cutoff. */
int oops_clusters_enumerate(const double* coords, int32_t nspin, double cutoff, int32_t max_order,
                            oops_clusters** out);
/* The same from a connection graph (cDepthFirstPathTracing on a given
 * connection matrix, SpinClusterAlgorithm.h:68-82): vertex i's neighbours are
 * neighbors[offsets[i] .. offsets[i+1]); the graph is made symmetric and its
 * diagonal ignored, as getConnectionMatrix builds it. */
int oops_clusters_enumerate_graph(int32_t nspin, const int64_t* offsets, const int32_t* neighbors,
                                  int32_t max_order, oops_clusters** out);
int oops_clusters_count(const oops_clusters* C, int32_t order, int64_t* count);
/* Spin indices, count x (order+1), clusters in the reference's set order. */
int oops_clusters_get(const oops_clusters* C, int32_t order, int32_t* idx);
/* Parent positions (in order-1) of each cluster: CSR offsets [count+1] + list. */
int oops_clusters_parents(const oops_clusters* C, int32_t order, int64_t* offsets, int32_t* list);
int oops_clusters_destroy(oops_clusters* C);

/* --- single-sample CCE (a pure bath state) ---------------------------------
 * Bath states of SingleSampleCCE::prepare_bath_state
 * (source/app/single_sample_cce.cpp:35-46): srand(seed), then spin i in
 * spin-list order is in basis state rand() % 2 (0: m = +1/2, 1: m = -1/2),
 * with glibc's rand() restated (the process's own rand() state is untouched). */
int oops_bath_states(uint32_t seed, int32_t nspin, int32_t* states);

/* Hilbert-space conditional Hamiltonian of one cluster with the bath's mean
 * field (SingleSampleCCE::create_spin_hamiltonian, single_sample_cce.cpp:75-93:
 * dipolar + Zeeman + NV field + sum over the bath spins outside the cluster of
 * their dipolar field in their basis states).  coords: all nspin bath spins;
 * cluster: k ascending spin indices; which 0: H0 (NV state0), 1: H1 (state1);
 * psi (2^k complex, may be NULL): the cluster's product state
 * (create_cluster_state, :95-105). */
int oops_cce_single_sample_terms(const double* coords, int32_t nspin, const int32_t* cluster, int32_t k,
                                 const oops_cce_params* P, const int32_t* states, int32_t which, oops_terms** out,
                                 double* psi);

/* Batch of every cluster of C under the single-sample evolution
 * (SingleSampleCCE::cluster_evolution, single_sample_cce.cpp:48-73):
 * psi_1(t) = riffle(H0, H1, N) segments and psi_2(t) = riffle(H1, H0, N)
 * segments applied to psi (Pulse_Interval fractions, npulse = 1: Hahn echo),
 * L_c(t) = Re <psi_1(t), psi_2(t)> (calc_observables, :120-128) at the times
 * t[0..nt).  States of dimension <= 16 run on the dense path, larger ones on
 * the flip-mask Lanczos kernel (precision 1, or 4 for FP64).  Results through
 * oops_hahn_batch_results; launch with oops_echo_batch_run. */
int oops_cce_single_sample_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C,
                                        const oops_cce_params* P, const int32_t* states, int32_t npulse, int32_t nt,
                                        const double* t, int32_t m, int32_t klim, double tol, int32_t precision,
                                        oops_hahn_batch** out);

/* Compiled tables of one cluster's generator (which 0: lvA, 1: -lvB^dagger,
 * 2: lvB, 3: -lvA^dagger) from the host compiler (device 0) or the device term
 * builder (device 1, clusters of 1-4 spins).  The batch creators above build
 * on the device for unpolarised baths of clusters of at most four spins
 * (environment OOPS_TERM_BUILDER=host keeps the host compiler).  counts[6]:
 * local operators, ELL entries, widest row, one-spin and two-spin flip-mask
 * slots, Hermitian flag; locs: 6 int32 per local (bodies, spin, second spin,
 * row width, first entry, rows); val (2 doubles per entry), off, eslot per
 * entry (up to max_loc / max_ent); masks: 32 int32 (9 one-spin then 12
 * two-spin slot masks).  No reference counterpart: it exposes the layout
 * both builders produce so that they can be compared. */
int oops_cce_op_tables(oops_ctx* ctx, const double* coords, int32_t k, const oops_cce_params* P, int32_t which,
                       int32_t device, int32_t max_loc, int32_t max_ent, int32_t* counts, int32_t* locs, double* val,
                       int32_t* off, uint8_t* eslot, int32_t* masks);

/* CCE reduction of L (all clusters' rows, orders concatenated, nt columns):
 * tilde (same shape), final_result / each_order (nt x max_order, row-major),
 * and the sign / log-magnitude of final_result.  Outputs may be NULL. */
int oops_cce_reduce(const oops_clusters* C, int32_t nt, const double* L, double* tilde, double* final_result,
                    double* each_order, double* log_abs_final, int32_t* sign_final);

/* Build the Hahn-echo pair of sweeps of every selected cluster and create the
 * device batch.  Clusters are numbered globally, order 0 first and set order
 * within an order; sel (nsel ascending global indices) picks a subset, or
 * NULL for every cluster.  Items keep that order. */
int oops_cce_hahn_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C, const oops_cce_params* P,
                               int64_t nsel, const int64_t* sel, int32_t m, int32_t klim, double tol,
                               int32_t precision, oops_hahn_batch** out);

/* Pulse sequence on every cluster (oops_echo_item semantics, same time grid
 * t[0..nt) for all clusters: sweep times t/2 for npulse = 1, times t for
 * CPMG-N); operators built on the host's cores.  Items are ordered as the
 * clusters; the batch runs with oops_echo_batch_run / _launch. */
int oops_cce_echo_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C, const oops_cce_params* P,
                               int32_t npulse, int32_t nt, const double* t, int32_t m, int32_t klim, double tol,
                               int32_t precision, oops_hahn_batch** out);

#ifdef __cplusplus
}
#endif

#endif /* OOPS_CCE_H */
