// field.h -- the deformed-field re-bake of a hex lattice on the GPU
// (field.hip; not part of the C ABI).
//
// Reference (paths relative to the Talus tree):
//   src/talus/deformable/deformedfield/deformedfield.cc   bakeDeformedField
//   src/talus/deformable/contact/contact.cc:886-1425      deformedProbe,
//                                                         placeDeformedProbe,
//                                                         probeDeformed
//   src/talus/field/sampledsdf/sampledsdf.cc:234-320, 432-480  the probe
//
// The probe's tables (boundary and crack faces, per-cell bounds, centres and
// centre Jacobian inverses, the uniform grid of cells with their padded
// bounds as a CSR) are O(cells) and built on the host; the per-sample work
// (grid lookup, Newton inversion of each candidate cell, the rest field's
// trilinear read) runs one device thread per sample of the field.
#pragma once

#include <hip/hip_runtime_api.h>

#include "../../include/talus_mi.h"
#include "hexlattice.h"

namespace tmi {

// bakeDeformedField of lattice body `b` at positions x (3 per node, FP64)
// under the model [R | t] (3x4 row-major); current->distances rewritten in
// place; returns false when current is not on rest's lattice
bool bake_deformed_field(const LatticeBody &b, const double *x, const double *model, const talus_mi_sdf &rest,
                         bool carve, talus_mi_sdf &current, talus_mi_sample_box &changed, hipStream_t stream);

}  // namespace tmi
