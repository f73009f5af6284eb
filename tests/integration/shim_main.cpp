// Caller of the reference-side wrappers in cl_mi355x.cpp, shaped like
// cl_sr_system.cpp:168-216: one InputInfo of two expressions over a shared
// X|y, evaluated through evaluate_gpu_simple_wrapper and
// evaluate_gpu_mse_wrapper; prints one MSE per line.
#include <cstdio>
#include <string>
#include <vector>

#include "utils/utils.h"

void evaluate_gpu_mse_wrapper(InputInfo &input_info, float ***all_vars, std::vector<float> &mses, void *ctx,
                              bool upload_X, bool clear_cache, RunStats &stats);
void evaluate_gpu_simple_wrapper(InputInfo &input_info, float ***all_vars, std::vector<float> &mses, void *ctx,
                                 bool upload_X, RunStats &stats);
void *create_gpu_context();
void destroy_gpu_context(void *ctx);
void *create_gpu_simple_context();
void destroy_gpu_simple_context(void *ctx);

int main()
{
    const int nv = 2, N = 1000;
    std::vector<std::vector<float>> cols(nv + 1, std::vector<float>(N));
    for (int d = 0; d < N; ++d) {
        cols[0][d] = 1.0f + 0.004f * d;
        cols[1][d] = 2.0f - 0.001f * d;
        cols[2][d] = cols[0][d] * cols[1][d];  // y = x0 * x1
    }
    float *shared[nv + 1];
    for (int v = 0; v <= nv; ++v) shared[v] = cols[v].data();
    float **all_vars[2] = {shared, shared};
    // x0 * x1 (exact) and x0 + 1.5
    int t0[] = {3, -1, -1}, t1[] = {1, -1, 0};
    double v0[] = {0.0, 0.0, 1.0}, v1[] = {0.0, 0.0, 1.5};
    int nvars[] = {nv, nv}, ndps[] = {N, N}, ntok[] = {3, 3};
    int *toks[] = {t0, t1};
    double *vals[] = {v0, v1};
    std::string files[2];
    InputInfo in{};
    in.num_exprs = 2;
    in.num_vars = nvars;
    in.num_dps = ndps;
    in.num_tokens = ntok;
    in.tokens = toks;
    in.values = vals;
    in.data_filenames = files;
    in.max_tokens = 3;
    in.max_num_dps = N;
    in.max_num_features = nv + 1;
    std::vector<float> mses;
    RunStats rs;
    void *simple = create_gpu_simple_context();
    evaluate_gpu_simple_wrapper(in, all_vars, mses, simple, true, rs);
    for (float m : mses) std::printf("simple %.9g\n", m);
    destroy_gpu_simple_context(simple);
    void *cached = create_gpu_context();
    evaluate_gpu_mse_wrapper(in, all_vars, mses, cached, true, false, rs);
    evaluate_gpu_mse_wrapper(in, all_vars, mses, cached, false, false, rs);
    for (float m : mses) std::printf("cached %.9g\n", m);
    destroy_gpu_context(cached);
    return 0;
}
