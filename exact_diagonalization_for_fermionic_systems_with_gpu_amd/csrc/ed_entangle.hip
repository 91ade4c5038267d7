// ed_entangle.hip -- entanglement entropy from particle-number-block reduced density
// matrices (placeholder entry points; the block kernels follow).
#include "ed_common.hpp"

using namespace ed;

extern "C" int ed_entanglement_entropy(int, int, int, int64_t, int, const void *, double *, int64_t *, void *)
{
    return fail(ED_EUNSUPPORTED, "ed_entanglement_entropy: not built yet");
}

extern "C" int ed_entanglement_spectrum(int, int, int, int64_t, int, const void *, double *, int64_t, int64_t *,
                                        void *)
{
    return fail(ED_EUNSUPPORTED, "ed_entanglement_spectrum: not built yet");
}
