// device_scratch.h -- the device lattice the drop-in's free functions use for
// host GaugeFields (observables of a host field, the heatbath, hot start).
//
// One FP64 libqcbflow handle per extent, created on first use.  It is held by
// an object that is never destroyed, so no HIP call runs during static
// destruction (the HIP runtime may already be gone then); a caller that wants
// the device memory back before the process ends calls
// release_device_scratch().
#pragma once
#include <vector>

#include "gauge/GaugeField.h"
#include "qcbflow.h"

namespace qcbcompat {
// The scratch handle for geo's extents, holding field's interior links.
qcbf_flow* upload(const GaugeField& field, const GeometryCB& geo);
// The scratch handle for geo's extents (its field as left by the last call).
qcbf_flow* scratch(const GeometryCB& geo);
// field's interior (and periodic halos) from the scratch handle.
void download(GaugeField& field, const GeometryCB& geo);
// Destroys the scratch handle, if any.
void release_device_scratch();
// Interior site index ((t*L + z)*L + y)*L + x, x..t = 0..L-1, of an
// extended-box site (halo sites map to their periodic images).
long interior_site(const GeometryCB& geo, size_t site);
}  // namespace qcbcompat
