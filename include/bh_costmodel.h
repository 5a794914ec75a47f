/* bh_costmodel — the learned kernel cost model of the hip vector engine.
 *
 * A small multilayer perceptron (features -> hidden -> hidden -> 1, ReLU) that
 * predicts the natural log of a fused block's device time in microseconds from
 * the block's features (bohrium_amd/csrc/ve_hip/features.hpp).  Inference,
 * training (mean squared error on the log time) and Adam updates run on the GPU
 * in hand-written HIP kernels; parameters and data are fp32.
 *
 * The reference has no counterpart: its fuser scores merges with an analytic
 * weight, the bytes of arrays that become temporaries (core/jitk/graph.cpp:136-170).
 * With `fuser_list = learned, ...` the engine scores a merge of blocks a and b by
 * the predicted time saved, cost(a) + cost(b) - cost(a u b), and stops merging
 * when no merge is predicted to save time.  Without an installed model the
 * learned fuser falls back to the analytic greedy fuser.
 *
 * All functions return 0 on success, non-zero on failure (message in
 * bhcm_last_error()).  Host arrays are row-major float32.
 */
#ifndef BH_COSTMODEL_H
#define BH_COSTMODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bhcm_model bhcm_model;

const char *bhcm_last_error(void);

/* A model with `nfeat` inputs and two hidden layers of `hidden` units (hidden <= 64),
 * parameters drawn with the given seed (He initialisation). */
int bhcm_create(int nfeat, int hidden, uint64_t seed, bhcm_model **out);
int bhcm_destroy(bhcm_model *m);
/* Parameter count and layout: W1[nfeat][hidden], b1[hidden], W2[hidden][hidden], b2[hidden], w3[hidden], b3. */
int64_t bhcm_num_params(const bhcm_model *m);
int bhcm_get_params(bhcm_model *m, float *out);
int bhcm_set_params(bhcm_model *m, const float *in);
/* Features are standardised on the device as (x - mean) * inv_std before the first layer. */
int bhcm_set_normalization(bhcm_model *m, const float *mean, const float *inv_std);
/* y[i] = model(x[i]) for n rows of nfeat features. */
int bhcm_predict(bhcm_model *m, const float *x, int64_t n, float *y);
/* Loss = mean((model(x) - t)^2) and its gradient with respect to every parameter (bhcm_num_params values). */
int bhcm_gradient(bhcm_model *m, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on (x, t); loss[k] after step k when non-NULL. */
int bhcm_train(bhcm_model *m, const float *x, const float *t, int64_t n, int steps, float lr, float *loss);
/* Make `m` the model the engine's learned fuser uses (NULL: none).  The caller keeps ownership. */
int bhcm_install(bhcm_model *m);
/* Make `m` the launch-shape model of the engine's autotuner (NULL: none): it predicts a kernel's
 * log(us) from the 16 block features followed by 3 launch features (log2 of the largest launch's
 * workgroups, log2 of the elements a thread issues together, log2 of the elements per thread).
 * With autotune_mode = model the tuner launches the candidate with the lowest prediction and runs
 * no trials; in every mode the tuner reports the prediction of each candidate it measures.  The
 * caller keeps ownership. */
int bhcm_install_launch(bhcm_model *m);

#ifdef __cplusplus
}
#endif

#endif
