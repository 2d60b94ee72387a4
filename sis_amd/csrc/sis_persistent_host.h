// Host drivers of the persistent and batched step forms.
#pragma once
int choose_mode(sis_gpu_ctx *c) { (void)c; return 1; }
void persistent_steps(sis_gpu_ctx *c, int64_t nsteps) { for (int64_t s = 0; s < nsteps; s++) step_phases(c); }
void batched_steps(sis_gpu_ctx *c, int64_t nsteps) { for (int64_t s = 0; s < nsteps; s++) step_phases(c); }
