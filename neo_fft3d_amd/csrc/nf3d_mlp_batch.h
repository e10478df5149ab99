// Library-internal entry of the learned-PSD network used by the engine (not
// part of the C ABI).
#pragma once

struct nf3d_mlp;

// Log-power features of nseg compact-spectrum arrays (specs[s]: rows[s]
// blocks of n_in coefficients) into consecutive rows of x_out (BF16), in one
// launch on `stream`.  nseg is 1..8.  Returns 0, or -1 on bad arguments or a
// launch failure.
int nf3d_mlp_features_batch(nf3d_mlp* m, const void* const* specs, const int* rows, int nseg, void* x_out,
                            void* stream);
