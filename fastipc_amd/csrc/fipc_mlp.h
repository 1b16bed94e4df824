// fipc_mlp.h -- device-side interface of the warm-start network used by the
// context entry points (the C-ABI of the network is in include/fastipc.h).
#pragma once
#include <hip/hip_runtime.h>

struct fipc_mlp;
// input / output widths
int mlp_dims(const fipc_mlp* m, int* n_in, int* n_out);
// forward of n rows of device input X on stream s; *Y = device output
// (n x n_out) in the network's own buffer, valid until its next call
int mlp_forward_dev(fipc_mlp* m, hipStream_t s, int n, const float* X, const float** Y);
// append n (feature, target) rows from device memory to the dataset
int mlp_dataset_append_dev(fipc_mlp* m, hipStream_t s, int n, const float* X, const float* T);
// the network's own stream (synchronised before another stream uses it)
hipStream_t mlp_stream(fipc_mlp* m);
