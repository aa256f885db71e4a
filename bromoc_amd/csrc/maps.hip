// maps.hip -- grid files of the reference (READPHI, READRFPAR) and the REPWALLS layering of
// the repulsion map.  The readers are host code: they parse the files' Fortran records
// and hand back host arrays that bromoc_create uploads.  REPWALLS runs on the device: every
// layer is one stencil pass over the map (HBM-bound, a few bytes per node), where the
// reference loops over the nodes on the CPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bromoc_gpu.h"

namespace {

// One Fortran unformatted sequential record: 4-byte length markers around the payload; a
// negative leading marker means the record continues in the next subrecord.
bool read_record(FILE *f, std::vector<unsigned char> &out) {
    out.clear();
    for (;;) {
        int32_t head = 0, tail = 0;
        if (std::fread(&head, 4, 1, f) != 1) return false;
        const bool more = head < 0;
        const int64_t n = head < 0 ? -(int64_t)head : head;
        const size_t at = out.size();
        out.resize(at + (size_t)n);
        if (n && std::fread(out.data() + at, 1, (size_t)n, f) != (size_t)n) return false;
        if (std::fread(&tail, 4, 1, f) != 1) return false;
        if ((tail < 0 ? -(int64_t)tail : tail) != n) return false;
        if (!more) return true;
    }
}

struct MapFile {
    bromoc_map_header h{};
    std::vector<unsigned char> values;   // the float32 record(s) after the two headers
};

// Records 1 and 2 (readphi.f90:43-44, rfpar.f90:435-436); nvalues float32 records follow.
int read_map_file(const char *path, int nvalues, MapFile &m, std::vector<std::vector<float>> *vals) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return BROMOC_E_ARG;
    std::vector<unsigned char> r;
    int st = BROMOC_OK;
    if (!read_record(f, r) || r.size() != 3 * 4 + 4 * 8) st = BROMOC_E_ARG;
    if (!st) {
        std::memcpy(&m.h.nclx, r.data(), 4);
        std::memcpy(&m.h.ncly, r.data() + 4, 4);
        std::memcpy(&m.h.nclz, r.data() + 8, 4);
        std::memcpy(&m.h.dcel, r.data() + 12, 8);
        std::memcpy(&m.h.xbcen, r.data() + 20, 8);
        std::memcpy(&m.h.ybcen, r.data() + 28, 8);
        std::memcpy(&m.h.zbcen, r.data() + 36, 8);
        if (m.h.nclx < 1 || m.h.ncly < 1 || m.h.nclz < 1 || !(m.h.dcel > 0.0)) st = BROMOC_E_ARG;
    }
    if (!st && (!read_record(f, r) || r.size() != 6 * 8)) st = BROMOC_E_ARG;
    if (!st) std::memcpy(m.h.rec2, r.data(), 48);
    const size_t ncel3 = (size_t)m.h.nclx * m.h.ncly * m.h.nclz;
    if (vals) vals->clear();
    for (int k = 0; !st && k < nvalues; k++) {
        if (!read_record(f, r) || r.size() != 4 * ncel3) { st = BROMOC_E_ARG; break; }
        if (vals) {
            vals->emplace_back(ncel3);
            std::memcpy(vals->back().data(), r.data(), r.size());
        }
    }
    std::fclose(f);
    return st;
}

// The crop of readphi.f90:98-134 (and rfpar.f90:509-536): node ranges x1..x2 (1-based) that
// cover the system box plus one node, the cropped geometry, and the new centre.
struct Crop {
    int x1, x2, y1, y2, z1, z2;
};

Crop crop_to_box(bromoc_map_header &h, const double *box) {
    const double dcel = h.dcel;
    const double tranx = 0.5 * (h.nclx - 1) * dcel, trany = 0.5 * (h.ncly - 1) * dcel,
                 tranz = 0.5 * (h.nclz - 1) * dcel;
    Crop c;
    c.x1 = (int)((box[0] + tranx - h.xbcen) / dcel);
    c.y1 = (int)((box[2] + trany - h.ybcen) / dcel);
    c.z1 = (int)((box[4] + tranz - h.zbcen) / dcel);
    c.x2 = (int)((box[1] + tranx - h.xbcen) / dcel) + 2;
    c.y2 = (int)((box[3] + trany - h.ybcen) / dcel) + 2;
    c.z2 = (int)((box[5] + tranz - h.zbcen) / dcel) + 2;
    if (c.x1 < 1) c.x1 = 1;
    if (c.y1 < 1) c.y1 = 1;
    if (c.z1 < 1) c.z1 = 1;
    if (c.x2 > h.nclx) c.x2 = h.nclx;
    if (c.y2 > h.ncly) c.y2 = h.ncly;
    if (c.z2 > h.nclz) c.z2 = h.nclz;
    const double xmmn = (c.x1 - 1) * dcel - tranx + h.xbcen;
    const double ymmn = (c.y1 - 1) * dcel - trany + h.ybcen;
    const double zmmn = (c.z1 - 1) * dcel - tranz + h.zbcen;
    const double xmmm = (c.x2 - 1) * dcel - tranx + h.xbcen;
    const double ymmm = (c.y2 - 1) * dcel - trany + h.ybcen;
    const double zmmm = (c.z2 - 1) * dcel - tranz + h.zbcen;
    h.xbcen = 0.5 * (xmmm + xmmn);
    h.ybcen = 0.5 * (ymmm + ymmn);
    h.zbcen = 0.5 * (zmmm + zmmn);
    return c;
}

// Copy the crop of one grid (nxo x nyo x nzo, x-major) in the reference's loop order.
template <typename T, typename F>
void copy_crop(const Crop &c, int nyo, int nzo, const float *src, T *dst, F conv) {
    size_t posn = 0;
    for (int i = c.x1; i <= c.x2; i++)
        for (int j = c.y1; j <= c.y2; j++)
            for (int k = c.z1; k <= c.z2; k++)
                dst[posn++] = conv(src[(size_t)(i - 1) * nyo * nzo + (size_t)(j - 1) * nzo + (k - 1)]);
}

// One REPWALLS layer n (repwalls.f90:32-58): out starts as a copy of in; a node of the
// interior (1+n..ncl-n, 1-based) holding n becomes n+1 when its 27-node neighbourhood
// sums to 27n.  count: interior nodes that hold n.
__global__ void k_repwall(const int8_t *__restrict__ in, int8_t *__restrict__ out, int nx, int ny,
                          int nz, int n, unsigned int *count) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncel3 = (int64_t)nx * ny * nz;
    if (idx >= ncel3) return;
    const int c = (int)(idx % nz);
    const int b = (int)((idx / nz) % ny);
    const int a = (int)(idx / ((int64_t)ny * nz));
    // 0-based interior: n .. ncl-1-n
    if (a < n || a > nx - 1 - n || b < n || b > ny - 1 - n || c < n || c > nz - 1 - n) return;
    if (in[idx] != n) return;
    atomicAdd(count, 1u);
    int sum = 0;
    for (int i = a - 1; i <= a + 1; i++)
        for (int j = b - 1; j <= b + 1; j++)
            for (int k = c - 1; k <= c + 1; k++)
                sum += in[((int64_t)i * ny + j) * nz + k];
    if (sum == 27 * n) out[idx] = (int8_t)(n + 1);
}

}  // namespace

extern "C" {

int bromoc_read_map(const char *path, int32_t kind, int32_t adjust, const double *box,
                    bromoc_map_header *hdr, void *data, int64_t capacity) {
    if (!path || !hdr || (kind != BROMOC_MAP_PHIX && kind != BROMOC_MAP_PHIV) ||
        (adjust && !box))
        return BROMOC_E_ARG;
    MapFile m;
    std::vector<std::vector<float>> vals;
    int st = read_map_file(path, data ? 1 : 0, m, data ? &vals : nullptr);
    if (st) return st;
    const int nxo = m.h.nclx, nyo = m.h.ncly, nzo = m.h.nclz;
    bromoc_map_header h = m.h;
    Crop c{1, nxo, 1, nyo, 1, nzo};
    if (adjust) {
        c = crop_to_box(h, box);
        h.nclx = c.x2 - c.x1 + 1;
        h.ncly = c.y2 - c.y1 + 1;
        h.nclz = c.z2 - c.z1 + 1;
        if (h.nclx < 1 || h.ncly < 1 || h.nclz < 1) return BROMOC_E_ARG;
    }
    *hdr = h;
    if (!data) return BROMOC_OK;
    const int64_t ncel3 = (int64_t)h.nclx * h.ncly * h.nclz;
    if (capacity < ncel3) return BROMOC_E_CAPACITY;
    const float *src = vals[0].data();
    if (kind == BROMOC_MAP_PHIX)
        copy_crop(c, nyo, nzo, src, (float *)data, [](float v) { return v; });
    else   // readphi.f90:319-327: accessible where the file value is nonzero
        copy_crop(c, nyo, nzo, src, (int8_t *)data,
                  [](float v) { return (int8_t)(v != 0.0f ? 0 : 1); });
    return BROMOC_OK;
}

int bromoc_read_rfpar(int32_t nfiles, const char *const *paths, int32_t adjust,
                      const double *box, bromoc_map_header *hdr, double *radion,
                      float *gsrfen, float *greff, int64_t capacity) {
    if (nfiles < 1 || !paths || !hdr || !radion || (adjust && !box) ||
        (gsrfen != nullptr) != (greff != nullptr))
        return BROMOC_E_ARG;
    const bool want = gsrfen != nullptr;
    std::vector<MapFile> files(nfiles);
    std::vector<std::vector<std::vector<float>>> vals(nfiles);
    for (int f = 0; f < nfiles; f++) {
        if (!paths[f]) return BROMOC_E_ARG;
        int st = read_map_file(paths[f], want ? 2 : 0, files[f], want ? &vals[f] : nullptr);
        if (st) return st;
        const bromoc_map_header &a = files[0].h, &b = files[f].h;
        // rfpar.f90:455-463: every file on the same lattice
        if (a.nclx != b.nclx || a.ncly != b.ncly || a.nclz != b.nclz || a.dcel != b.dcel ||
            a.xbcen != b.xbcen || a.ybcen != b.ybcen || a.zbcen != b.zbcen)
            return BROMOC_E_ARG;
        radion[f] = b.rec2[0];
    }
    const int nxo = files[0].h.nclx, nyo = files[0].h.ncly, nzo = files[0].h.nclz;
    bromoc_map_header h = files[0].h;
    Crop c{1, nxo, 1, nyo, 1, nzo};
    if (adjust) {
        c = crop_to_box(h, box);
        h.nclx = c.x2 - c.x1 + 1;
        h.ncly = c.y2 - c.y1 + 1;
        h.nclz = c.z2 - c.z1 + 1;
    }
    *hdr = h;
    if (!want) return BROMOC_OK;
    const int64_t ncel3 = (int64_t)h.nclx * h.ncly * h.nclz;
    if (capacity < ncel3 * nfiles) return BROMOC_E_CAPACITY;
    for (int f = 0; f < nfiles; f++) {
        copy_crop(c, nyo, nzo, vals[f][0].data(), gsrfen + f * ncel3, [](float v) { return v; });
        copy_crop(c, nyo, nzo, vals[f][1].data(), greff + f * ncel3, [](float v) { return v; });
    }
    return BROMOC_OK;
}

int bromoc_repwalls(int32_t nclx, int32_t ncly, int32_t nclz, int8_t *phiv, int32_t walls,
                    int device) {
    if (!phiv || nclx < 1 || ncly < 1 || nclz < 1 || walls < 0 || walls > 126)
        return BROMOC_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return BROMOC_E_NODEVICE;
    if (walls == 0) return BROMOC_OK;
    if (hipSetDevice(device) != hipSuccess) return BROMOC_E_HIP;
    const size_t n = (size_t)nclx * ncly * nclz;
    int8_t *buf[2] = {nullptr, nullptr};
    unsigned int *cnt = nullptr;
    int st = BROMOC_OK;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BROMOC_E_HIP;
    if (hipMalloc(&buf[0], n) != hipSuccess || hipMalloc(&buf[1], n) != hipSuccess ||
        hipMalloc(&cnt, sizeof(unsigned int)) != hipSuccess)
        st = BROMOC_E_HIP;
    if (!st && hipMemcpy(buf[0], phiv, n, hipMemcpyHostToDevice) != hipSuccess) st = BROMOC_E_HIP;
    int cur = 0;
    const unsigned int zero = 0;
    for (int layer = 1; !st && layer <= walls; layer++) {
        unsigned int ck = 0;
        if (hipMemcpyAsync(buf[cur ^ 1], buf[cur], n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(cnt, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess) {
            st = BROMOC_E_HIP;
            break;
        }
        const unsigned int blocks = (unsigned int)((n + 255) / 256);
        hipLaunchKernelGGL(k_repwall, dim3(blocks), dim3(256), 0, s, buf[cur], buf[cur ^ 1],
                           nclx, ncly, nclz, layer, cnt);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&ck, cnt, sizeof(ck), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            st = BROMOC_E_HIP;
            break;
        }
        cur ^= 1;
        if (ck == 0) break;   // repwalls.f90:57
    }
    if (!st && hipMemcpy(phiv, buf[cur], n, hipMemcpyDeviceToHost) != hipSuccess) st = BROMOC_E_HIP;
    for (int b = 0; b < 2; b++)
        if (buf[b]) (void)hipFree(buf[b]);
    if (cnt) (void)hipFree(cnt);
    (void)hipStreamDestroy(s);
    return st;
}

}  // extern "C"
