// Learned subgrid closure: device state of the training side.
#pragma once

#include <hip/hip_runtime.h>

#include "bl_internal.hpp"

namespace blgpu {

constexpr int kMlpIn = 10;      // g_ij / |g| (9) and y / delta99
constexpr int kMlpHidden = 64;  // three hidden layers of 64, SiLU
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + 2 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpHidden + 1;
// dataset row: features, target c, weight, Delta^2 |S|, S_ij (6), tau_ij (6), pad
constexpr int kMlpRow = 28;
constexpr int kRowTarget = kMlpIn, kRowWeight = kMlpIn + 1, kRowScale = kMlpIn + 2, kRowS = kMlpIn + 3,
              kRowTau = kMlpIn + 9;

struct Mlp {
    long long cap, n;   // dataset capacity and fill (rows)
    int bmax;           // largest training batch
    int nslice;         // partial-gradient slices of bmax
    long long step;     // Adam steps taken
    hipStream_t stream;
    float *rows;        // dataset [cap][kMlpRow]
    float *theta, *adam_m, *adam_v, *g, *part;
    float *Xb, *tb, *wb, *Z, *D, *y, *dy, *sq, *scal;
};

void mlp_init(Mlp &m, long long capacity, int batch_max);
void mlp_free(Mlp &m);
void mlp_load_batch(Mlp &m, const float *x, const float *t, int n);
void mlp_fwd_bwd(Mlp &m, int B, bool backward);
void mlp_train(Mlp &m, int steps, int B, float lr, unsigned long long seed, float *loss_hist);
// bl_closure.hip
void closure_samples(Handle &fine, Handle &coarse, Mlp &m, int fx, int fy, int fz, double delta99, double smin);
void closure_install(Handle &h, Mlp &m, double delta99, double nsigma);
void closure_fallback(Handle &h, long long *cells, long long *fallback);
void closure_infer_rows(Mlp &m, const float *x_host, long long n, float *y_host);

}  // namespace blgpu

struct bl_mlp : blgpu::Mlp {};
