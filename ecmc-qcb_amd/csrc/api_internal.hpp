// api_internal.hpp -- helpers shared by the single-lattice and ensemble ABIs.
#pragma once
#include <hip/hip_runtime.h>

#include "qcbflow.h"

namespace qcb {
// Precision (as qcb::prec), current link field and stream of a handle.
int field_view(qcbf_flow* h, int* prec, const void** field, hipStream_t* stream);
// Records msg as the error of a call made without a handle (what
// qcbf_last_error(NULL) returns on this thread) and returns code.
int fail_without_handle(int code, const char* msg);
}  // namespace qcb
