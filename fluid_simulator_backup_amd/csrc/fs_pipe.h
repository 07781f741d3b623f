// fs_pipe.h -- the oil/water pipe mode's solver (fs_create_pipe), host side.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/fs_abi.h"

struct FsPipe;

// Failures of the pipe solver: an FS_ERR_* code and a message.
struct FsPipeError {
    int code;
    std::string msg;
};

FsPipe* fs_pipe_new(const fs_params& P, const fs_pipe_params& Q, const fs_grid& g, int device);
void fs_pipe_delete(FsPipe* s);
void fs_pipe_initial(FsPipe* s);
void fs_pipe_consts(FsPipe* s, double* out);
void fs_pipe_set(FsPipe* s, const double* phi, const double* u, const double* v, const double* w, const double* p);
void fs_pipe_get(FsPipe* s, double* phi, double* u, double* v, double* w, double* p);
// one step (frame: the level set's ComputeCurvatures first); FS_OK or
// FS_ERR_NOT_CONVERGED (fields updated) -- other failures throw
int fs_pipe_advance(FsPipe* s, double dt, fs_step_stats* st, bool frame);
int fs_pipe_run_phase(FsPipe* s, int phase, double dt, fs_step_stats* st);
double fs_pipe_cfl(FsPipe* s, double cfl);
void fs_pipe_array(FsPipe* s, int which, double* dst);
void fs_pipe_async(FsPipe* s, bool async);
void fs_pipe_sync(FsPipe* s);
hipStream_t fs_pipe_stream(FsPipe* s);
// in-loop samples of the PCG loop kernels (0 k_pcg_spmv_dir, 1 k_pcg_update2)
void fs_pipe_sampling(FsPipe* s, bool on);
void fs_pipe_samples(FsPipe* s, int kernel, double* avg_ms, double* avg_bytes, long long* samples, long long* launches);
