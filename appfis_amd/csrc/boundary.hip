// On-device boundary fill and single-rank halo exchange.
//
// appfis_set_boundary       <- Grid::SetBoundary       (reference Appfis/Grid.hpp:558-678)
// appfis_set_boundary_copy  <- Grid::SetBoundaryCopy   (reference Appfis/Grid.hpp:682-830)
// appfis_exchange           <- Communication::Exchange / ExchangeGPU at world size 1
//                              (reference Appfis/Communication.hpp:613-748 with the face
//                              pack/unpack of Runtime.hpp:917-1173): a periodic axis maps
//                              each rank to itself, so pack -> Sendrecv -> unpack collapses
//                              to one on-device copy per axis; no host staging.
#include "common.hpp"

using namespace appfis;

namespace {

// axis of a face and whether it is the high side, following APPFIS::DIRECTION.
int face_axis(int ndim, int face, int* high)
{
    *high = 0;
    if (ndim == 1) {
        if (face == 0) return 0;
        if (face == 1) { *high = 1; return 0; }
        return -1;
    }
    if (ndim == 2) {
        switch (face) {
        case 0: return 0;
        case 1: *high = 1; return 0;
        case 2: return 1;
        case 3: *high = 1; return 1;
        default: return -1;
        }
    }
    switch (face) {
    case 0: return 0;
    case 1: *high = 1; return 0;
    case 4: return 1;
    case 5: *high = 1; return 1;
    case 3: return 2;
    case 2: *high = 1; return 2;
    default: return -1;
    }
}

// A face is addressed as (g, a, b): g the layer, (a, b) the two other axes in
// increasing axis order.  coords() turns that into (i, j, k).
struct FaceMap {
    int ax, oa, ob;  // face axis and the other two axes
    int d[3];
    __device__ size_t offset(int layer, int a, int b) const
    {
        int c[3];
        c[ax] = layer;
        c[oa] = a;
        c[ob] = b;
        return (size_t)c[0] + (size_t)d[0] * ((size_t)c[1] + (size_t)d[1] * (size_t)c[2]);
    }
};

FaceMap make_face(int ax, const int* d)
{
    FaceMap f;
    f.ax = ax;
    f.oa = ax == 0 ? 1 : 0;
    f.ob = ax == 2 ? 1 : 2;
    for (int a = 0; a < 3; a++) f.d[a] = d[a];
    return f;
}

template <typename E>
__global__ void k_set_face(E* data, FaceMap f, int first_layer, int ghost, E value)
{
    const int na = f.d[f.oa], nb = f.d[f.ob];
    size_t n = (size_t)ghost * na * nb;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int a = (int)(t % (size_t)na);
    size_t r = t / (size_t)na;
    int b = (int)(r % (size_t)nb);
    int g = (int)(r / (size_t)nb);
    data[f.offset(first_layer + g, a, b)] = value;
}

// One thread per (a, b) line; walks g from ghost-1 down to 0 like the reference,
// so the in-line dependency of the high face (ghost > 1) is reproduced exactly.
template <typename E>
__global__ void k_copy_face(E* data, FaceMap f, int high, int ghost)
{
    const int na = f.d[f.oa], nb = f.d[f.ob];
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)na * nb) return;
    int a = (int)(t % (size_t)na);
    int b = (int)(t / (size_t)na);
    const int dim = f.d[f.ax];
    for (int g = ghost - 1; g >= 0; g--) {
        int dst = high ? dim - ghost + g : g;
        int src = high ? dim - ghost + g - 1 : g + 1;
        data[f.offset(dst, a, b)] = data[f.offset(src, a, b)];
    }
}

// Periodic self-exchange of one axis: high ghost layers <- interior layers [g, 2g),
// low ghost layers <- interior layers [dim-2g, dim-g).  The four layer ranges are
// disjoint when the interior holds at least 2*ghost layers (checked on the host).
template <typename E>
__global__ void k_wrap_axis(E* data, FaceMap f, int ghost)
{
    const int na = f.d[f.oa], nb = f.d[f.ob];
    size_t n = (size_t)ghost * na * nb;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int a = (int)(t % (size_t)na);
    size_t r = t / (size_t)na;
    int b = (int)(r % (size_t)nb);
    int g = (int)(r / (size_t)nb);
    const int dim = f.d[f.ax];
    E lo = data[f.offset(ghost + g, a, b)];
    E hi = data[f.offset(dim - 2 * ghost + g, a, b)];
    data[f.offset(dim - ghost + g, a, b)] = lo;
    data[f.offset(g, a, b)] = hi;
}

template <typename E>
int set_faces(appfis_ctx* ctx, void* data, int ndim, const int* d, int ghost, int face, const void* value)
{
    E v;
    memcpy(&v, value, sizeof(E));
    int f0 = face < 0 ? 0 : face, f1 = face < 0 ? 2 * ndim : face + 1;
    for (int fc = f0; fc < f1; fc++) {
        int high;
        int ax = face_axis(ndim, fc, &high);
        if (ax < 0) return invalid("face %d is not valid for a %dD array", fc, ndim);
        FaceMap f = make_face(ax, d);
        size_t n = (size_t)ghost * d[f.oa] * d[f.ob];
        if (n == 0) continue;
        int first = high ? d[ax] - ghost : 0;
        hipLaunchKernelGGL(k_set_face<E>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (E*)data, f, first,
                           ghost, v);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

template <typename E>
int copy_faces(appfis_ctx* ctx, void* data, int ndim, const int* d, int ghost, int face)
{
    int f0 = face < 0 ? 0 : face, f1 = face < 0 ? 2 * ndim : face + 1;
    for (int fc = f0; fc < f1; fc++) {
        int high;
        int ax = face_axis(ndim, fc, &high);
        if (ax < 0) return invalid("face %d is not valid for a %dD array", fc, ndim);
        FaceMap f = make_face(ax, d);
        size_t n = (size_t)d[f.oa] * d[f.ob];
        hipLaunchKernelGGL(k_copy_face<E>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (E*)data, f, high,
                           ghost);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

template <typename E>
int wrap_axes(appfis_ctx* ctx, void* data, int ndim, const int* d, int ghost)
{
    for (int ax = 0; ax < ndim; ax++) {
        FaceMap f = make_face(ax, d);
        size_t n = (size_t)ghost * d[f.oa] * d[f.ob];
        hipLaunchKernelGGL(k_wrap_axis<E>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (E*)data, f, ghost);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

}  // namespace

extern "C" {

int appfis_set_boundary(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost, int face,
                        const void* value)
{
    APPFIS_REQUIRE(ctx && data && value, "NULL argument");
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(ghost >= 0, "ghost must be >= 0");
    APPFIS_REQUIRE(face >= -1 && face < 2 * ndim, "face %d is not valid for a %dD array", face, ndim);
    for (int a = 0; a < ndim; a++) APPFIS_REQUIRE(d[a] >= ghost, "ghost %d exceeds dims[%d] = %d", ghost, a, d[a]);
    if (ghost == 0) return APPFIS_OK;
    switch (dtype_size(dtype)) {
    case 1: return set_faces<uint8_t>(ctx, data, ndim, d, ghost, face, value);
    case 2: return set_faces<uint16_t>(ctx, data, ndim, d, ghost, face, value);
    case 4: return set_faces<uint32_t>(ctx, data, ndim, d, ghost, face, value);
    case 8: return set_faces<uint64_t>(ctx, data, ndim, d, ghost, face, value);
    default: return invalid("unsupported dtype %d", dtype);
    }
}

int appfis_set_boundary_copy(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost, int face)
{
    APPFIS_REQUIRE(ctx && data, "NULL argument");
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(ghost >= 0, "ghost must be >= 0");
    APPFIS_REQUIRE(face >= -1 && face < 2 * ndim, "face %d is not valid for a %dD array", face, ndim);
    for (int a = 0; a < ndim; a++)
        APPFIS_REQUIRE(d[a] >= ghost + 1, "dims[%d] = %d leaves no layer to copy from", a, d[a]);
    if (ghost == 0) return APPFIS_OK;
    switch (dtype_size(dtype)) {
    case 1: return copy_faces<uint8_t>(ctx, data, ndim, d, ghost, face);
    case 2: return copy_faces<uint16_t>(ctx, data, ndim, d, ghost, face);
    case 4: return copy_faces<uint32_t>(ctx, data, ndim, d, ghost, face);
    case 8: return copy_faces<uint64_t>(ctx, data, ndim, d, ghost, face);
    default: return invalid("unsupported dtype %d", dtype);
    }
}

int appfis_exchange(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost, int periodic)
{
    APPFIS_REQUIRE(ctx && data, "NULL argument");
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(ghost >= 0, "ghost must be >= 0");
    if (!periodic || ghost == 0) return APPFIS_OK;
    for (int a = 0; a < ndim; a++)
        APPFIS_REQUIRE(d[a] - 2 * ghost >= 2 * ghost,
                       "subgrid extent %d along axis %d is below 2*ghost (Common.hpp:203-214)", d[a] - 2 * ghost, a);
    switch (dtype_size(dtype)) {
    case 1: return wrap_axes<uint8_t>(ctx, data, ndim, d, ghost);
    case 2: return wrap_axes<uint16_t>(ctx, data, ndim, d, ghost);
    case 4: return wrap_axes<uint32_t>(ctx, data, ndim, d, ghost);
    case 8: return wrap_axes<uint64_t>(ctx, data, ndim, d, ghost);
    default: return invalid("unsupported dtype %d", dtype);
    }
}

}  // extern "C"
