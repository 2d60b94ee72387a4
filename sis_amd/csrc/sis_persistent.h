// Persistent and batched step kernels (filled in below).
#pragma once
#include "sis_kernels.h"
