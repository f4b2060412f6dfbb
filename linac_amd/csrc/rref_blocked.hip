// Blocked right-looking Gauss-Jordan over GF(p) (placeholder: not yet built).
#include "gf_common.hpp"

namespace linac {

template <class F>
int blocked_dispatch(typename F::T*, int64_t, int64_t, int64_t, const F&, hipStream_t, const ElimWorkspace&,
                     DeviceContext*, bool, Profiler*) {
    return NOT_APPLICABLE;
}

template int blocked_dispatch<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);
template int blocked_dispatch<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);
template int blocked_dispatch<FieldGF2>(uint32_t*, int64_t, int64_t, int64_t, const FieldGF2&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);

}  // namespace linac
