/* cudimot_amd.h -- C-ABI of the MI355X voxelwise fitting engine.
 *
 * Drop-in boundary for cudimot's fitting path.  Each entry point replaces one
 * method object of the reference's C++ engine (constructor + run()), takes
 * plain device pointers and sizes, returns 0 on success or a negative status
 * with a message in cdm_last_error(), and never exits the process (the
 * reference prints and calls exit(-1): checkcudacalls.h:19-26, Model.cu:97).
 *
 * Buffers (device memory, caller-owned, FP32, voxel-major as the reference):
 *   meas     [nvox][nmeas]            dMRI_Data.cu:137 (meas_host[vox*nmeas+m])
 *   params   [nvox][nparams]          Parameters.cu:72,81 (in/out)
 *   cfp      [nmeas][cfp_tsize]       Parameters.cu:119 (CFP_host[i+j*CFP_Tsize])
 *   fixp     [nvox][fixp_tsize]       Parameters.cu:186
 *   samples  [nvox][nparams][nsamp]   MCMC.cu:392,621
 *   tau      [nvox][nsamp]            MCMC.cu:624
 * `stream` is a hipStream_t (NULL: the null stream); calls are asynchronous
 * on it.  A model handle is not safe for concurrent use from several host
 * threads.  Its LM work-queue counter, MCMC chain state and GridSearch
 * scratch (small volumes) belong to the handle: such a call on another
 * stream than the handle's previous one waits on the device for that call
 * to finish (outside stream capture; a capturing caller orders its streams
 * itself).  Use one handle per stream for concurrent fits.
 */
#ifndef CUDIMOT_AMD_H
#define CUDIMOT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cdm_model cdm_model;

enum cdm_method { CDM_GRIDSEARCH = 0, CDM_LEVMAR = 1, CDM_MCMC = 2, CDM_PREDICTED = 3, CDM_BIC_AIC = 4 };

/* Library version (major*10000 + minor*100 + patch). */
int cdm_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* cdm_last_error(void);

/* Comma-separated names of the model plug-ins compiled into this library.
 * Returns the full length needed (excluding the terminator). */
int cdm_list_models(char* buf, int buflen);

/* Open a compiled plug-in and parse its modelpriors file.
 * Replaces Model<T>::Model + Modelparser (Model.cu:16-299, 395-403); the
 * reference selects the model at compile time (Makefile:3-5, one binary per
 * model) and reads <binary>_priors or --priors (Model.cu:20-25).
 * priors_file: path, or NULL for the plug-in's own modelpriors. */
int cdm_model_open(const char* name, const char* priors_file, cdm_model** out);
void cdm_model_close(cdm_model* m);

/* NPARAMS and the summed CFP / FixP row sizes (Model.cu:248-260). */
int cdm_model_info(const cdm_model* m, int* nparams, int* cfp_tsize, int* fixp_tsize);

/* Number and row sizes of the common fixed parameters (CFP, per
 * measurement) and of the per-voxel fixed parameters (FixP): the plug-in's
 * MODEL::CFP_size / FixP_size (template/modelparameters.h:17-21; read by
 * Model.cu:248-260).  Size arrays may be NULL. */
int cdm_model_sizes(const cdm_model* m, int* ncfp, int* cfp_sizes, int* nfixp, int* fixp_sizes);

/* Parsed specification (arrays of nparams; any pointer may be NULL).
 * Bounds and prior arguments are the parsed doubles; the kernels use them
 * rounded to float, as the reference's __constant__ float copies
 * (GridSearch.cu:271-280, Levenberg_Marquardt.cu:643-660, MCMC.cu:723-750). */
int cdm_model_spec(const cdm_model* m, int* bound_types, double* bmin, double* bmax, int* prior_types,
                   double* prior_a, double* prior_b, int* fixed, double* p_init, int* init_provided);

/* --fixed=i,j,... (Model.cu:279-298). */
int cdm_model_set_fixed(cdm_model* m, const char* list);

/* --gridSearch file (Model.cu:302-393).  is_text != 0: `src` is the file
 * text itself.  Outputs the number of combinations and grid parameters. */
int cdm_model_load_grid(cdm_model* m, const char* src, int is_text, int* ncombs, int* ngrid_params);
/* Grid read-back: grid_params [ngrid_params], grid [ncombs][ngrid_params]. */
int cdm_model_grid(const cdm_model* m, int* grid_params, double* grid);

/* Voxel-group size (lanes per voxel: 1, 4, 16 or 64) used by a method;
 * 0 restores the automatic choice: the tuned or built-in size for the
 * measurement count, widened for launches over few voxels (nvox x lanes
 * below ~131k lane slots) so that the grid still fills the GPU.  get_group
 * reports the size a launch of nvox voxels would use (nvox <= 0: no
 * widening). */
int cdm_model_set_group(cdm_model* m, int method, int lanes);
int cdm_model_get_group(const cdm_model* m, int method, int nmeas, int nvox);

/* Register budget of the LM and MCMC kernels, in resident 256-thread blocks
 * per CU the compiled variant allows for (1, 2 or 4: the compiler caps
 * registers, spilling beyond them, for more waves in flight); 0 restores the
 * automatic choice (tuned, else 1).  The Rician MCMC kernel has budget 1. */
int cdm_model_set_occupancy(cdm_model* m, int method, int blocks);
int cdm_model_get_occupancy(const cdm_model* m, int method, int nmeas);

/* Autotuned launch table (no counterpart in the reference, whose launch
 * shape is fixed: one warp per voxel, 64 voxels per block, MCMC.cu:780-784).
 * Lines "<model> <nmeas> <method> <lanes> [ms [blocks]]", method one of gridsearch,
 * levmar, mcmc, predicted, bic_aic; entries of other models are skipped.
 * A method without an explicit set_group uses the entry of the nearest
 * measurement count.  cdm_model_open loads <libdir>/../tuning.txt when it
 * exists.  path NULL clears the table.  Returns the entries kept. */
int cdm_model_load_tuning(cdm_model* m, const char* path);

/* MCMC jump blocks: split each MCMC phase into launches of at most
 * `jumps_per_launch` jumps (0: one launch per phase); use_graph != 0
 * records the launches of a phase into a HIP graph, kept on the handle and
 * replayed by later calls with the same arguments (buffers, sizes, seed,
 * bounds, launch shape); other arguments record a new graph. */
int cdm_model_set_mcmc_blocks(cdm_model* m, int jumps_per_launch, int use_graph);

/* Graphs recorded (captures) and replayed from the handle's cache (no
 * counterpart in the reference; for tests and launch tuning). */
int cdm_model_graph_stats(const cdm_model* m, long long* captures, long long* replays);

/* GridSearch<T>::run (GridSearch.cu:291-317; kernel :97-248).  Updates the
 * grid parameters of `params` to the combination of least squared error
 * (the FP64 argmin, earliest combination on ties).  kept (nullable,
 * device int[nvox]) receives the kept combination index, -1 if none.
 * Volumes too small to fill the GPU search slices of the combinations in
 * parallel, with the same result (scratch allocated on the handle at the
 * first such call of a larger size). */
int cdm_gridsearch(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
                   const float* fixp, int* kept, void* stream);

/* Levenberg_Marquardt<T>::run (Levenberg_Marquardt.cu:665-704; kernel
 * :409-623).  max_iters = --iterLevMar (successful iterations);
 * marquardt = !--no_Marquardt.  iters (int[nvox]) and cost (float[nvox])
 * are optional device outputs. */
int cdm_levmar(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
               const float* fixp, int max_iters, int marquardt, int* iters, float* cost, void* stream);

/* MCMC<T>::MCMC + run (MCMC.cu:690-848): burn-in (nburn jumps, at least 1)
 * then recording (njumps, a sample every sampleevery jumps into samples;
 * nsamples = njumps / sampleevery).  Proposal widths adapt every
 * sampleevery jumps (the reference ignores --upe, MCMC.cu:711).
 * rician: --rician (tau_samples required).  seed/voxel_offset address the
 * Philox streams: draws depend only on (seed, voxel_offset + voxel, jump). */
int cdm_mcmc(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
             const float* fixp, int nburn, int njumps, int sampleevery, int update_proposal, int rician,
             unsigned long long seed, long long voxel_offset, float* samples, float* tau_samples, void* stream);

/* cdm_mcmc without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and draws as cdm_mcmc (bit-identical samples and
 * parameters); the chain state (and, for models whose per-measurement cache
 * rows do not fit in LDS, those 2 x nmeas rows per voxel) lives in a
 * library-owned static device scratch of 64 MiB per GPU, held until the
 * process ends, instead of the handle's allocation.  A volume needing more
 * (NODDI_Bingham: 256 bytes of chain state per voxel, 262,144 voxels) is
 * refused with an error: split it and pass each part's voxel_offset, which
 * keeps the random streams.  Every call
 * on a GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * streams, other handles, other host threads) race.  No jump-block graphs
 * (cdm_model_set_mcmc_blocks' use_graph is ignored; jump blocks still
 * apply). */
int cdm_mcmc_noalloc(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
                     const float* fixp, int nburn, int njumps, int sampleevery, int update_proposal, int rician,
                     unsigned long long seed, long long voxel_offset, float* samples, float* tau_samples,
                     void* stream);

/* --debug=<voxel> (cudimotoptions.h:160-161).  The reference prints the
 * voxel's state from inside its kernels (GridSearch.cu:152-241,
 * Levenberg_Marquardt.cu:482-615, MCMC.cu:437-600); here GridSearch, LM
 * and MCMC write records of voxel `voxel` (the index within each call's
 * volume, as the reference's index within a data part) to a device buffer
 * of max_records records per method, which cdm_model_debug_trace copies out
 * after the call.  voxel < 0 turns it off.  A record is 4 + nparams doubles:
 *   GridSearch (one record): kept combination (enumeration index), its
 *     cost, valid combinations, 0, the parameters;
 *   LM (record 0 the start, record k after attempt k): k, cost, lambda
 *     after the attempt, 1 if accepted, the parameters;
 *   MCMC (record j after jump j, burn-in and recording jumps counted
 *     together): j, energy, tau (Rician), acceptances since the last
 *     proposal update, the parameters. */
int cdm_model_set_debug(cdm_model* m, int voxel, int max_records);

/* The records of the last call of `method` (CDM_GRIDSEARCH, CDM_LEVMAR,
 * CDM_MCMC) into out (up to max_records of them; out may be null);
 * *nrecords receives the number written by the call (at most the buffer's
 * max_records).  Waits for `stream`. */
int cdm_model_debug_trace(cdm_model* m, int method, double* out, int max_records, int* nrecords, void* stream);

/* getPredictedSignal<T>::run (getPredictedSignal.cu:98-118): signal of the
 * mean sample, out [nvox][nmeas]. */
int cdm_predicted_signal(cdm_model* m, int nvox, int nmeas, int nsamples, const float* samples, const float* cfp,
                         const float* fixp, float* out, void* stream);

/* BIC_AIC<T>::run (BIC_AIC.cu:212-243): BIC and AIC of the mean sample,
 * Rician form when rician != 0 (the reference uses it only with --rician
 * and --runMCMC, BIC_AIC.cu:208). */
int cdm_bic_aic(cdm_model* m, int nvox, int nmeas, int nsamples, int rician, const float* meas,
                const float* samples, const float* tau_samples, const float* cfp, const float* fixp, float* bic,
                float* aic, void* stream);

/* ------------------------------------------------------------------------
 * Learned initialiser.  No counterpart in the reference: it stands in for
 * GridSearch (GridSearch.cu:291-317) as the source of LM starting points.
 * A fully connected network n_in -> width (x depth, GELU) -> n_out, with
 * BF16 MFMA products and FP32 accumulation for inference and training.
 * Inputs are a voxel's measurements divided by a per-voxel scale (S0).
 * The output head maps network outputs to model parameters, one entry per
 * predicted parameter (or parameter pair, for an axis):
 *   CDM_HEAD_SIGMOID        p = lo + (hi - lo) s(o)                 1 output
 *   CDM_HEAD_SIGMOID_LOG    p = exp(log lo + (log hi - log lo) s(o)) 1 output
 *   CDM_HEAD_SIGMOID_RATIO  p = (lo + (hi - lo) s(o)) P[param2]       1 output
 *   CDM_HEAD_AXIS           (P[param], P[param2]) = (theta, phi) of the
 *                           output vector (o0, o1, o2)               3 outputs
 * with s the logistic function; decoded values are clamped to [vmin, vmax]
 * (axis entries excepted).  Training loss: weight x soft-target cross entropy
 * per sigmoid entry, weight x (1 - (v.n)^2 / |v|^2) per axis. */
typedef struct cdm_mlp cdm_mlp;

enum cdm_head_kind { CDM_HEAD_SIGMOID = 0, CDM_HEAD_SIGMOID_LOG = 1, CDM_HEAD_SIGMOID_RATIO = 2, CDM_HEAD_AXIS = 3 };

typedef struct cdm_head_entry {
  int kind;    /* cdm_head_kind */
  int param;   /* parameter index (theta for an axis) */
  int param2;  /* ratio reference, or phi for an axis; -1 otherwise */
  float lo, hi;
  float weight;     /* loss weight */
  float vmin, vmax; /* clamp of the decoded value */
} cdm_head_entry;

/* Create a network with LeCun-uniform weights from `seed`; head is a host
 * array of nhead entries (decoded in order). */
int cdm_mlp_create(int n_in, int width, int depth, const cdm_head_entry* head, int nhead, unsigned long long seed,
                   cdm_mlp** out);
void cdm_mlp_destroy(cdm_mlp* m);
int cdm_mlp_info(const cdm_mlp* m, int* n_in, int* width, int* depth, int* n_out, long long* nweights);
/* Host copy of the FP32 master weights, per layer W[out][in] then b[out]. */
int cdm_mlp_get_weights(const cdm_mlp* m, float* w);
/* Replace the weights (resets the optimiser state). */
int cdm_mlp_set_weights(cdm_mlp* m, const float* w);

/* Raw network outputs for n voxels: out [n][n_out] (device).  nmeas must be
 * the network's input width; scale (device float[n]) may be NULL (1). */
int cdm_mlp_forward(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, float* out, void* stream);

/* Learned starting point: writes the head's parameters into params
 * [n][nparams] (device); parameters without a head entry are left as they
 * are (the caller fills them, e.g. with P_init). */
int cdm_mlp_init_params(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, int nparams,
                        float* params, void* stream);

/* One forward + backward pass over n (multiple of 32) voxels of given
 * measurements without updating: mean loss and its gradient with respect to
 * every weight (host arrays, get_weights layout).  params_true [n][nparams]
 * (device) are the targets.  For gradient checks. */
int cdm_mlp_loss_grad(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, int nparams,
                      const float* params_true, float* loss, float* grad, void* stream);

/* Train with Adam (cosine-decayed learning rate) on n noise-free signals
 * clean [n][nmeas] (device) of known parameters params_true [n][nparams];
 * every epoch draws fresh Rician noise, sigma = scale / SNR with SNR uniform
 * in [snr_lo, snr_hi] per sample (Philox streams under `seed`).  batch is a
 * multiple of 32.  epoch_loss (host, [epochs], nullable) receives each
 * epoch's mean loss. */
int cdm_mlp_train(cdm_mlp* m, int n, int nmeas, const float* clean, const float* scale, int nparams,
                  const float* params_true, float snr_lo, float snr_hi, int epochs, int batch, float lr,
                  unsigned long long seed, float* epoch_loss, void* stream);

/* use_graph != 0 (the default): cdm_mlp_train records one training step into
 * a HIP graph and replays it once per step, the step's changing arguments
 * (batch offset, noise epoch, learning rate, Adam's bias corrections) read
 * from a device table; 0: every step's kernels are launched one by one.  The
 * same kernels and arguments either way: the weights are identical. */
int cdm_mlp_set_train_graph(cdm_mlp* m, int use_graph);

/* ------------------------------------------------------------------------
 * Batched persistent launch.  No counterpart in the reference, which runs one
 * binary per model over its voxel sub-parts in turn (cudimot.cc:81-120).
 * Volumes of several models (each item: its own model handle, measurement
 * count and buffers, with the layouts above) are fitted by one call per
 * method.  Items too small to fill the GPU on their own share ONE
 * persistent launch: workgroups take voxel tiles from a work queue,
 * heaviest items first, until all are done.  An item whose voxels alone
 * fill the GPU several times over (8 blocks of its voxel group per CU) runs
 * through its model's own kernels at its tuned voxel group and register
 * budget instead; the persistent kernel has one register allocation, the
 * largest of its models'.  Results and draws are the same either way.  The
 * models must be compiled together into a batch unit (python -m
 * cudimot_amd.build --batch NAME=A,B,...); each item uses its model
 * handle's bounds, priors, fixed parameters and tuned voxel-group size. */
typedef struct cdm_batch_item {
  cdm_model* model; /* a model handle used by no other item of the batch */
  int nvox, nmeas;
  const float* meas;
  float* params;
  const float* cfp;
  const float* fixp;
  int* iters;   /* LM outputs (nullable) */
  float* cost;
  float* samples;     /* MCMC outputs: [nvox][nparams][njumps / sampleevery] */
  float* tau_samples; /* MCMC with rician != 0: [nvox][njumps / sampleevery] */
  long long voxel_offset; /* MCMC random-stream offset */
} cdm_batch_item;

/* "name:ModelA,ModelB;name2:..." of the compiled batch units.  Returns the
 * full length needed. */
int cdm_batch_units(char* buf, int buflen);

/* Levenberg_Marquardt<T>::run over every item.  The batched calls are
 * asynchronous on `stream`; the persistent launch's item tables and
 * work-queue counters live in a per-process workspace (one per method and
 * device, grown as needed), so a batched call first waits on the host for
 * the previous call of the same method on that device to finish. */
int cdm_batch_levmar(int nitems, const cdm_batch_item* items, int max_iters, int marquardt, void* stream);

/* MCMC<T>::run over every item: one launch per jump block (jumps_per_launch
 * jumps, 0: a whole phase), each over all the persistent launch's items;
 * use_graph != 0 records a phase's launches into a HIP graph, replayed by
 * later calls of the same shape (the workspace's graphs,
 * cdm_batch_graph_stats; for the items on their own kernels, the handle's
 * graphs, cdm_model_graph_stats).  Draws are those of
 * cdm_mcmc with the same seed and voxel offsets.  rician != 0: the Rician
 * likelihood (MCMC.cu:799-847) for every item, each with its tau_samples. */
int cdm_batch_mcmc(int nitems, const cdm_batch_item* items, int nburn, int njumps, int sampleevery,
                   int update_proposal, int rician, unsigned long long seed, int jumps_per_launch, int use_graph,
                   void* stream);

/* Items whose voxels fill at least blocks_per_cu blocks of their voxel
 * group per CU run through their model's own kernels (default 8); 0 sends
 * every item through the persistent launch.  Process-wide; no counterpart in
 * the reference. */
int cdm_batch_set_fill(int blocks_per_cu);

/* Graphs recorded and replayed by cdm_batch_mcmc in this process (no
 * counterpart in the reference). */
int cdm_batch_graph_stats(long long* captures, long long* replays);

/* Orientation initialisation from a tensor fit (the reference pipeline's
 * dtifit V1/V2 -> cart2spherical -> initialise_Psi steps: utils/Run_dtifit.sh:36-46,
 * utils/cart2spherical.cc:17-69, utils/initialise_Psi.cc:18-106).  dti
 * [n][stride] (device) holds the tensor (Dxx, Dxy, Dxz, Dyy, Dyz, Dzz) at
 * column `first` (DiffusionTensor plug-in: first = 1).  out [n][6] (device):
 * theta, phi of the principal eigenvector, psi of the second eigenvector in
 * the Bingham frame, and the eigenvalues in descending order. */
int cdm_tensor_frame(int n, const float* dti, int stride, int first, float* out, void* stream);

/* ------------------------------------------------------------------------
 * Derived maps of the <Model>_finish.sh steps, from MCMC samples
 * [nvox][nparams][nsamples] on the device (or parameters, nsamples = 1). */

/* Mean of each parameter's samples: mean [nvox][nparams] (fslmaths -Tmean
 * of each Param_k_samples volume, e.g. NODDI_Watson_finish.sh:30-32). */
int cdm_sample_mean(int nvox, int nparams, int nsamples, const float* samples, float* mean, void* stream);

/* Mean orientation: principal eigenvector of the mean dyadic tensor of the
 * sampled directions, dyads [nvox][3] (sign: z >= 0), and dispersion
 * [nvox] = 1 - its eigenvalue (nullable).  psi < 0: fibre directions
 * (theta, phi) as FSL's make_dyadic_vectors (Ball_1_Stick_finish.sh:31);
 * psi >= 0: Bingham fanning directions (utils/getFanningOrientation.cc:39-76). */
int cdm_dyads(int nvox, int nparams, int nsamples, const float* samples, int theta, int phi, int psi, float* dyads,
              float* dispersion, void* stream);

/* NODDI dispersion maps from the kappa (and beta) samples.  beta < 0
 * (Watson): maps [nvox][1] = OD = 2/pi atan(1/mean kappa)
 * (NODDI_Watson_finish.sh:35).  beta >= 0 (Bingham, k2 = kappa - beta):
 * maps [nvox][5] = mean k2, ODIp, ODIs, ODItot, DA
 * (NODDI_Bingham_finish.sh:32-56). */
int cdm_noddi_maps(int nvox, int nparams, int nsamples, const float* samples, int kappa, int beta, float* maps,
                   void* stream);

/* Host-side data preparation for --rician: per voxel, replace non-positive
 * signals by the smallest positive one when some are positive
 * (dMRI_Data.cu:14-31).  meas is HOST memory [nvox][nmeas]. */
int cdm_rician_clean(int nvox, int nmeas, float* meas);

/* Block until all work queued on `stream` has finished; reports any
 * asynchronous device error (the reference's sync_check). */
int cdm_sync(void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CUDIMOT_AMD_H */
