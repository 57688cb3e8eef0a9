// Step-driver side of the DFT record (monitors.hip): the batched form records a step into
// a snapshot slot and adds a full batch to the accumulators; dft_flush adds what is
// pending.  Bit-identical to recording each step with mp_probe_dft_accumulate.
#pragma once
#include "common.h"

namespace mp {
int dft_batch_size();
int dft_record_batched(const mp_yee *g, const mp_plane *planes, int32_t nplanes, const double *omega_bins,
                       double *phase, int32_t nb, double t, double dt);
int dft_flush(const mp_yee *g, const mp_plane *planes, int32_t nplanes, int32_t nb, double dt);
// Whether the batch holds as many steps as it can; the next record flushes it first.
bool dft_batch_full();
// The current device's address of the fixed scratch of mp_probe_shot_reduce_noalloc
// (resolved once per device; mp_init resolves it for the device it selects).
int noalloc_scratch(double **out);
}  // namespace mp
