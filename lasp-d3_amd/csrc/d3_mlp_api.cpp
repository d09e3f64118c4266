// C ABI of the stand-in network's kernels on caller-owned device buffers (include/d3x_sur.h):
// packing, inference, training gradients (energy and force losses), Adam.  Used by the trainer
// (lasp-d3_amd/surrogate/train.py) and by tests; the engine calls the launchers directly.
#include <hip/hip_runtime_api.h>

#include "../../include/d3x.h"
#include "../../include/d3x_sur.h"
#include "d3_host.hpp"
#include "d3_sur.hpp"

namespace {

using namespace d3::sur;

static_assert(D3X_SUR_NPARAMS == NPARAMS, "parameter count");
static_assert(D3X_SUR_NFEAT == NFEAT, "feature count");

// Status of the launches just enqueued.
int launched(const char *where) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    d3::set_last_error(std::string(where) + ": " + hipGetErrorString(e));
    return (int)(D3X_ERR_GENERIC | D3X_ERR_DEVICE);
}

int bad(const char *where, const char *what) {
    d3::set_last_error(std::string(where) + ": " + what);
    return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
}

}  // namespace

extern "C" {

size_t d3x_sur_packed_bytes(void) { return PACKED_BYTES; }

size_t d3x_sur_train_bytes(uint64_t rows) {
    return (size_t)((rows + 15) / 16) * 16 * (TRAIN_FORCE_ROW_BYTES > TRAIN_ROW_BYTES ? TRAIN_FORCE_ROW_BYTES
                                                                                    : TRAIN_ROW_BYTES);
}

int d3x_sur_pack(const float *params, void *packed, void *stream) {
    if (!params || !packed) return bad("d3x_sur_pack", "NULL buffer");
    launch_pack(params, packed, (hipStream_t)stream);
    return launched("d3x_sur_pack");
}

int d3x_sur_infer(const float *params, const void *packed, uint64_t n, const float *X, const int32_t *cls, float *E,
                  float *G, void *stream) {
    if (n == 0) return 0;
    if (!params || !packed || !X || !cls) return bad("d3x_sur_infer", "NULL buffer");
    if (n > (uint64_t)INT32_MAX) return bad("d3x_sur_infer", "too many rows");
    launch_infer(params, packed, (int)n, X, cls, E, G, (hipStream_t)stream);
    return launched("d3x_sur_infer");
}

int d3x_sur_train_grad(const float *params, const void *packed, int cls, uint64_t n, const float *X, const float *y,
                       const float *w, float *E, void *work, float *grad, float *loss, void *stream) {
    if (n == 0) return 0;
    if (!params || !packed || !X || !y || !w || !E || !work || !grad || !loss)
        return bad("d3x_sur_train_grad", "NULL buffer");
    if (cls < 0 || cls >= NCLASS) return bad("d3x_sur_train_grad", "class out of range");
    if (n > (uint64_t)INT32_MAX) return bad("d3x_sur_train_grad", "too many rows");
    launch_train(params, packed, cls, (int)n, X, y, w, E, work, grad, loss, (hipStream_t)stream);
    return launched("d3x_sur_train_grad");
}

int d3x_sur_train_grad_noalloc(const float *params, const void *packed, int cls, uint64_t n, const float *X,
                               const float *y, const float *w, float *E, float *grad, float *loss, void *stream) {
    static_assert(D3X_SUR_NOALLOC_MAX_ROWS == STATIC_TRAIN_ROWS, "static scratch rows");
    if (n == 0) return 0;
    if (!params || !packed || !X || !y || !w || !E || !grad || !loss)
        return bad("d3x_sur_train_grad_noalloc", "NULL buffer");
    if (cls < 0 || cls >= NCLASS) return bad("d3x_sur_train_grad_noalloc", "class out of range");
    if (n > (uint64_t)STATIC_TRAIN_ROWS)
        return bad("d3x_sur_train_grad_noalloc", "more rows than the static scratch holds (at most 16384 per call)");
    void *work = nullptr;
    const hipError_t e = static_train_scratch(&work);
    if (e != hipSuccess) {
        d3::set_last_error(std::string("d3x_sur_train_grad_noalloc: static scratch: ") + hipGetErrorString(e));
        return (int)(D3X_ERR_GENERIC | D3X_ERR_DEVICE);
    }
    launch_train(params, packed, cls, (int)n, X, y, w, E, work, grad, loss, (hipStream_t)stream);
    return launched("d3x_sur_train_grad_noalloc");
}

int d3x_sur_train_force_grad(const float *params, const void *packed, int cls, uint64_t n, const float *X,
                             const float *U, const float *w, float *D, void *work, float *grad, void *stream) {
    if (n == 0) return 0;
    if (!params || !packed || !X || !U || !w || !D || !work || !grad)
        return bad("d3x_sur_train_force_grad", "NULL buffer");
    if (cls < 0 || cls >= NCLASS) return bad("d3x_sur_train_force_grad", "class out of range");
    if (n > (uint64_t)INT32_MAX) return bad("d3x_sur_train_force_grad", "too many rows");
    launch_train_force(params, packed, cls, (int)n, X, U, w, D, work, grad, (hipStream_t)stream);
    return launched("d3x_sur_train_force_grad");
}

int d3x_sur_tie_grad(float *grad, int blocks, void *stream) {
    if (!grad) return bad("d3x_sur_tie_grad", "NULL buffer");
    launch_tie(grad, blocks != 0 ? 1 : 0, (hipStream_t)stream);
    return launched("d3x_sur_tie_grad");
}

int d3x_sur_adam(float *params, float *grad, float *m, float *v, float lr, float beta1, float beta2, float eps,
                 int step, void *stream) {
    if (!params || !grad || !m || !v) return bad("d3x_sur_adam", "NULL buffer");
    if (step < 1) return bad("d3x_sur_adam", "step must be >= 1");
    launch_adam(params, grad, m, v, lr, beta1, beta2, eps, step, (hipStream_t)stream);
    return launched("d3x_sur_adam");
}

}  // extern "C"
