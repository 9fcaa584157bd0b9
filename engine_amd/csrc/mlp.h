// mlp.h -- sizes and parameter offsets of the learned rho MLP (internal).
#pragma once

#define MLP_NIN 6
#define MLP_NH 32
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_NH * MLP_NIN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_NH)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_NH * MLP_NH)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_NH)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_NH)
#define MLP_NPARAMS (MLP_OFF_B3 + 1)
#define MLP_NPARAMS_PAD 1316

// policy output range: rho = 10^clamp(y, MLP_LOG10_RHO_MIN, MLP_LOG10_RHO_MAX)
#define MLP_LOG10_RHO_MIN (-1.0f)
#define MLP_LOG10_RHO_MAX (4.0f)

// edge types of the feature vector
#define PDM_EDGE_PAIR 0
#define PDM_EDGE_CONTAINER 1
#define PDM_EDGE_RADIUS 2

struct paradmm_rho_mlp;

namespace pdm {
const float* mlp_device_params(paradmm_rho_mlp* m);
}
