// fieldio.hip -- the reference's field files (io.f) between disk and host or device memory.
//
// IOSCALAR (io.f:114-200) and IOVECTOR (io.f:8-108) use Fortran sequential unformatted files:
// every WRITE statement is one record, framed by its byte length as a 4-byte integer before
// and after the payload (the layout of ifort -r8 and of gfortran / flang -fdefault-real-8:
// INTEGER 4 bytes, REAL 8 bytes, little-endian). Record sequences:
//   DIR=1 (one file for the whole field, written by rank 0):
//     IOSCALAR: [NX NY NZG 1] + NZG records, each the plane P(1:NX,1:NY,K) +
//               [NSTEP] [TIME] [DTM1 GRAV]                                  (io.f:144-173)
//     IOVECTOR: [NX NY NZG 3] + NZG records, each the U, V and W planes of K (io.f:47-83)
//   DIR=2 (one file per rank, NAME//PROC(MYRANK)):
//     IOSCALAR: [NX NY NZ 1 TIME DTM1 GRAV] + NZ plane records               (io.f:186-191)
//     IOVECTOR: [NX NY NZ 3 TIME] + NZ records of U, V, W planes             (io.f:96-102)
// with NZG = MYSIZE*(NZ-2)+2. Rank r of MYSIZE reads the NZ planes from r*(NZ-2)+1 on: its
// slab with both ghost planes, as the reference's reads skip r*(NZ-2) records first
// (io.f:26-28, 125-129). The reference's IOVECTOR read expects TIME in the header record
// (io.f:24) that its DIR=1 write does not put there (io.f:47): both header forms are read.
// Unlike the reference, which reads with the caller's NX, NY unchecked, every header and
// record length is checked against the caller's dimensions before data is stored.
//
// A plane is contiguous in A(NX,NY,NZ), so a run of plane records is a strided block: the
// device paths move runs of records through two pinned staging buffers with one 2-D copy
// per field and run (source pitch = record length + 8 marker bytes), so that the file read
// or write of one run overlaps the copy of the other.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eddyproj.h"
#include "plan.h"

namespace eddy {
namespace {

constexpr size_t kStageBytes = 32u << 20;  // per staging buffer

int io_error(const std::string& path, const std::string& what) {
    set_error(path + ": " + what);
    return EDDY_ERR_IO;
}

struct File {
    int fd = -1;
    ~File() {
        if (fd >= 0) ::close(fd);
    }
};

bool read_at(int fd, void* dst, size_t n, off_t off) {
    char* p = static_cast<char*>(dst);
    while (n > 0) {
        const ssize_t r = ::pread(fd, p, n, off);
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += r;
    }
    return true;
}

bool write_all(int fd, const void* src, size_t n) {
    const char* p = static_cast<const char*>(src);
    while (n > 0) {
        const ssize_t r = ::write(fd, p, n);
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
    }
    return true;
}

// One record's payload at offset off: checks both markers, returns the payload length.
bool record_at(int fd, off_t off, int32_t& len) {
    int32_t tail = 0;
    if (!read_at(fd, &len, 4, off) || len < 0) return false;
    return read_at(fd, &tail, 4, off + 4 + len) && tail == len;
}

// The layout of a field file as found on disk.
struct Layout {
    eddy_field_info info{};
    off_t first_plane = 0;  // offset of the first plane record's leading marker
    size_t rec = 0;         // plane record payload bytes (nvar * NX * NY * 8)
};

int scan(const std::string& path, int fd, Layout& lay) {
    int32_t hlen = 0;
    if (!record_at(fd, 0, hlen) || hlen < 16 || hlen > 64)
        return io_error(path, "no header record of 4 integers (not a Fortran unformatted field file?)");
    std::vector<char> h((size_t)hlen);
    if (!read_at(fd, h.data(), h.size(), 4)) return io_error(path, "short header record");
    int32_t dims[4];
    std::memcpy(dims, h.data(), 16);
    eddy_field_info& in = lay.info;
    in = eddy_field_info{};
    in.nx = dims[0];
    in.ny = dims[1];
    in.nz = dims[2];
    in.nvar = dims[3];
    if (in.nx < 1 || in.ny < 1 || in.nz < 1 || (in.nvar != 1 && in.nvar != 3))
        return io_error(path, "header is not NX, NY, NZ, 1|3");
    const double* hd = reinterpret_cast<const double*>(h.data() + 16);
    if (in.nvar == 1 && hlen == 16 + 3 * 8) {  // IOSCALAR DIR=2
        in.per_rank = 1;
        in.has_time = 1;
        std::memcpy(&in.time, hd, 8);
        std::memcpy(&in.dtm1, hd + 1, 8);
        std::memcpy(&in.grav, hd + 2, 8);
    } else if (in.nvar == 3 && hlen == 16 + 8) {  // IOVECTOR with TIME in the header
        in.has_time = 1;
        std::memcpy(&in.time, hd, 8);
    } else if (hlen != 16) {
        return io_error(path, "header record of " + std::to_string(hlen) + " bytes matches no IOSCALAR/IOVECTOR layout");
    }
    lay.first_plane = 8 + hlen;
    lay.rec = (size_t)in.nvar * (size_t)in.nx * (size_t)in.ny * sizeof(double);
    if (lay.rec > 0x7fffffffu) return io_error(path, "plane record of 2 GiB or more");
    // first and last plane records, then the IOSCALAR DIR=1 trailer
    int32_t len = 0;
    const off_t step = (off_t)lay.rec + 8;
    for (off_t k : {(off_t)0, (off_t)in.nz - 1}) {
        if (!record_at(fd, lay.first_plane + k * step, len) || (size_t)len != lay.rec)
            return io_error(path, "plane record " + std::to_string(k + 1) + " missing or not " +
                                      std::to_string(lay.rec) + " bytes");
    }
    if (in.nvar == 1 && !in.per_rank) {
        off_t t = lay.first_plane + (off_t)in.nz * step;
        int32_t n1 = 0, n2 = 0, n3 = 0;
        if (!record_at(fd, t, n1) || n1 != 4 || !record_at(fd, t + 12, n2) || n2 != 8 ||
            !record_at(fd, t + 28, n3) || n3 != 16)
            return io_error(path, "IOSCALAR trailer [NSTEP] [TIME] [DTM1 GRAV] missing");
        int32_t nstep = 0;
        double dg[2];
        if (!read_at(fd, &nstep, 4, t + 4) || !read_at(fd, &in.time, 8, t + 16) || !read_at(fd, dg, 16, t + 32))
            return io_error(path, "short trailer");
        in.nstep = nstep;
        in.has_time = 1;
        in.dtm1 = dg[0];
        in.grav = dg[1];
    }
    return 0;
}

int open_scan(const std::string& path, File& f, Layout& lay) {
    f.fd = ::open(path.c_str(), O_RDONLY);
    if (f.fd < 0) return io_error(path, "cannot open for reading");
    return scan(path, f.fd, lay);
}

// The first file plane (0-based) of rank myrank's slab of nz planes, after checking the
// caller's dimensions against the file.
int slab_start(const std::string& path, const Layout& lay, int nx, int ny, int nz, int myrank, int mysize,
               int nvar, long& k0) {
    const eddy_field_info& in = lay.info;
    if (nx < 1 || ny < 1 || nz < 3 || mysize < 1 || myrank < 0 || myrank >= mysize) {
        set_error("bad dimensions or rank for a field read");
        return EDDY_ERR_ARG;
    }
    if (in.nvar != nvar) return io_error(path, nvar == 1 ? "not an IOSCALAR file" : "not an IOVECTOR file");
    if (in.nx != nx || in.ny != ny)
        return io_error(path, "file holds NX, NY = " + std::to_string(in.nx) + ", " + std::to_string(in.ny) +
                                  "; caller has " + std::to_string(nx) + ", " + std::to_string(ny));
    k0 = (long)myrank * (nz - 2);
    if (k0 + nz > in.nz)
        return io_error(path, "file has " + std::to_string(in.nz) + " planes; rank " + std::to_string(myrank) +
                                  " needs planes " + std::to_string(k0 + 1) + ".." + std::to_string(k0 + nz));
    if (mysize == 1 && in.nz != nz)
        return io_error(path, "file has " + std::to_string(in.nz) + " planes, caller NZ = " + std::to_string(nz));
    return 0;
}

// Planes [k0, k0 + nz) into host arrays (nvar of them), markers checked.
int read_host(const std::string& path, int fd, const Layout& lay, long k0, int nz, double* const* dst) {
    const size_t plane = lay.rec / (size_t)lay.info.nvar;
    const off_t step = (off_t)lay.rec + 8;
    for (int k = 0; k < nz; ++k) {
        const off_t off = lay.first_plane + (k0 + k) * step;
        int32_t len = 0;
        if (!record_at(fd, off, len) || (size_t)len != lay.rec)
            return io_error(path, "plane record " + std::to_string(k0 + k + 1) + " is damaged");
        for (int v = 0; v < lay.info.nvar; ++v)
            if (!read_at(fd, dst[v] + (size_t)k * (plane / sizeof(double)), plane, off + 4 + (off_t)(v * plane)))
                return io_error(path, "short plane record " + std::to_string(k0 + k + 1));
    }
    return 0;
}

struct Staging {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};  // a copy through buf[b] was enqueued and not yet waited for
    int wait(int b) {
        if (!pending[b]) return 0;
        pending[b] = false;
        const hipError_t e = hipEventSynchronize(ev[b]);
        return e == hipSuccess ? 0 : device_error(e, "hipEventSynchronize (field staging)");
    }
    int mark(int b, hipStream_t st) {
        const hipError_t e = hipEventRecord(ev[b], st);
        if (e != hipSuccess) return device_error(e, "hipEventRecord");
        pending[b] = true;
        return 0;
    }
    ~Staging() {  // an early error return leaves copies in flight: wait before freeing
        for (int b = 0; b < 2; ++b)
            if (pending[b]) (void)hipEventSynchronize(ev[b]);
        for (int b = 0; b < 2; ++b) {
            if (ev[b]) (void)hipEventDestroy(ev[b]);
            if (buf[b]) (void)hipHostFree(buf[b]);
        }
    }
    int init(size_t bytes) {
        for (int b = 0; b < 2; ++b) {
            hipError_t e = hipHostMalloc((void**)&buf[b], bytes, hipHostMallocDefault);
            if (e != hipSuccess) return device_error(e, "hipHostMalloc (field staging)");
            e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
            if (e != hipSuccess) return device_error(e, "hipEventCreate");
        }
        return 0;
    }
};

// Planes [k0, k0 + nz) into device arrays through the staging buffers.
int read_dev(const std::string& path, int fd, const Layout& lay, long k0, int nz, double* const* dst,
             hipStream_t st) {
    const size_t plane = lay.rec / (size_t)lay.info.nvar, step = lay.rec + 8;
    const int run = (int)std::max<size_t>(1, kStageBytes / step);
    Staging sg;
    int rc = sg.init(std::min<size_t>((size_t)run, (size_t)nz) * step);
    if (rc) return rc;
    for (int k = 0, b = 0; k < nz; k += run, b ^= 1) {
        const int n = std::min(run, nz - k);
        if ((rc = sg.wait(b))) return rc;  // the copy out of this buffer is done
        const off_t off = lay.first_plane + (off_t)(k0 + k) * (off_t)step;
        if (!read_at(fd, sg.buf[b], (size_t)n * step, off)) return io_error(path, "short read of plane records");
        for (int r = 0; r < n; ++r) {
            int32_t h, t;
            std::memcpy(&h, sg.buf[b] + (size_t)r * step, 4);
            std::memcpy(&t, sg.buf[b] + (size_t)r * step + 4 + lay.rec, 4);
            if ((size_t)h != lay.rec || t != h)
                return io_error(path, "plane record " + std::to_string(k0 + k + r + 1) + " is damaged");
        }
        for (int v = 0; v < lay.info.nvar; ++v) {
            const hipError_t e =
                hipMemcpy2DAsync(dst[v] + (size_t)k * (plane / sizeof(double)), plane, sg.buf[b] + 4 + (size_t)v * plane,
                                 step, plane, (size_t)n, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return device_error(e, "hipMemcpy2DAsync (field read)");
        }
        if ((rc = sg.mark(b, st))) return rc;
    }
    const hipError_t e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : device_error(e, "hipStreamSynchronize (field read)");
}

int write_record(int fd, const void* payload, int32_t len) {
    return (write_all(fd, &len, 4) && write_all(fd, payload, (size_t)len) && write_all(fd, &len, 4)) ? 0 : -1;
}

struct Header {
    int32_t dims[4];
    double extra[3];
    int nextra;
};

// Header, nz planes of nvar fields (host or device), then the IOSCALAR DIR=1 trailer.
int write_field(const std::string& path, const Header& hd, const double* const* src, int nvar, int nz, bool device,
                hipStream_t st, const int32_t* nstep, const double* trailer_time, const double* dtm1_grav) {
    const size_t plane = (size_t)hd.dims[0] * (size_t)hd.dims[1] * sizeof(double);
    const size_t rec = (size_t)nvar * plane, step = rec + 8;
    if (rec > 0x7fffffffu) return io_error(path, "plane record of 2 GiB or more");
    File f;
    f.fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (f.fd < 0) return io_error(path, "cannot open for writing");
    std::vector<char> h(16 + 8 * (size_t)hd.nextra);
    std::memcpy(h.data(), hd.dims, 16);
    std::memcpy(h.data() + 16, hd.extra, 8 * (size_t)hd.nextra);
    if (write_record(f.fd, h.data(), (int32_t)h.size())) return io_error(path, "write failed");
    const int32_t len = (int32_t)rec;
    if (!device) {
        for (int k = 0; k < nz; ++k) {
            if (!write_all(f.fd, &len, 4)) return io_error(path, "write failed");
            for (int v = 0; v < nvar; ++v)
                if (!write_all(f.fd, src[v] + (size_t)k * (plane / sizeof(double)), plane))
                    return io_error(path, "write failed");
            if (!write_all(f.fd, &len, 4)) return io_error(path, "write failed");
        }
    } else {
        const int run = (int)std::max<size_t>(1, kStageBytes / step);
        Staging sg;
        int rc = sg.init(std::min<size_t>((size_t)run, (size_t)nz) * step);
        if (rc) return rc;
        auto enqueue = [&](int k, int b) -> int {  // device planes [k, k + n) -> staging b
            const int n = std::min(run, nz - k);
            for (int v = 0; v < nvar; ++v) {
                hipError_t e = hipMemcpy2DAsync(sg.buf[b] + 4 + (size_t)v * plane, step,
                                                src[v] + (size_t)k * (plane / sizeof(double)), plane, plane, (size_t)n,
                                                hipMemcpyDeviceToHost, st);
                if (e != hipSuccess) return device_error(e, "hipMemcpy2DAsync (field write)");
            }
            return sg.mark(b, st);
        };
        if ((rc = enqueue(0, 0))) return rc;
        for (int k = 0, b = 0; k < nz; k += run, b ^= 1) {
            const int n = std::min(run, nz - k);
            if (k + run < nz && (rc = enqueue(k + run, b ^ 1))) return rc;  // next run in flight
            if ((rc = sg.wait(b))) return rc;
            for (int r = 0; r < n; ++r) {
                std::memcpy(sg.buf[b] + (size_t)r * step, &len, 4);
                std::memcpy(sg.buf[b] + (size_t)r * step + 4 + rec, &len, 4);
            }
            if (!write_all(f.fd, sg.buf[b], (size_t)n * step)) return io_error(path, "write failed");
        }
    }
    if (nstep) {
        if (write_record(f.fd, nstep, 4) || write_record(f.fd, trailer_time, 8) || write_record(f.fd, dtm1_grav, 16))
            return io_error(path, "write failed");
    }
    return 0;
}

int check_write_args(const char* path, const void* a, int nx, int ny, int nz) {
    if (!path || !a || nx < 1 || ny < 1 || nz < 1) {
        set_error("null path or field, or bad dimensions, for a field write");
        return EDDY_ERR_ARG;
    }
    return 0;
}

}  // namespace
}  // namespace eddy

using namespace eddy;

extern "C" {

int eddy_field_info_read(const char* path, eddy_field_info* info) {
    if (!path || !info) {
        set_error("null path or info");
        return EDDY_ERR_ARG;
    }
    File f;
    Layout lay;
    const int rc = open_scan(path, f, lay);
    if (rc == 0) *info = lay.info;
    return rc;
}

static int ioscalar_read_impl(const char* path, double* p, int nx, int ny, int nz, int myrank, int mysize,
                              eddy_field_info* info, bool device, void* stream) {
    if (!path || !p) {
        set_error("null path or field");
        return EDDY_ERR_ARG;
    }
    File f;
    Layout lay;
    int rc = open_scan(path, f, lay);
    long k0 = 0;
    if (!rc) rc = slab_start(path, lay, nx, ny, nz, myrank, mysize, 1, k0);
    if (rc) return rc;
    double* dst[1] = {p};
    rc = device ? read_dev(path, f.fd, lay, k0, nz, dst, (hipStream_t)stream) : read_host(path, f.fd, lay, k0, nz, dst);
    if (!rc && info) *info = lay.info;
    return rc;
}

int eddy_ioscalar_read(const char* path, double* p, int nx, int ny, int nz, int myrank, int mysize,
                       eddy_field_info* info) {
    return ioscalar_read_impl(path, p, nx, ny, nz, myrank, mysize, info, false, nullptr);
}

int eddy_ioscalar_read_dev(const char* path, double* d_p, int nx, int ny, int nz, int myrank, int mysize,
                           eddy_field_info* info, void* stream) {
    return ioscalar_read_impl(path, d_p, nx, ny, nz, myrank, mysize, info, true, stream);
}

static int iovector_read_impl(const char* path, double* u, double* v, double* w, int nx, int ny, int nz, int myrank,
                              int mysize, eddy_field_info* info, bool device, void* stream) {
    if (!path || !u || !v || !w) {
        set_error("null path or field");
        return EDDY_ERR_ARG;
    }
    File f;
    Layout lay;
    int rc = open_scan(path, f, lay);
    long k0 = 0;
    if (!rc) rc = slab_start(path, lay, nx, ny, nz, myrank, mysize, 3, k0);
    if (rc) return rc;
    double* dst[3] = {u, v, w};
    rc = device ? read_dev(path, f.fd, lay, k0, nz, dst, (hipStream_t)stream) : read_host(path, f.fd, lay, k0, nz, dst);
    if (!rc && info) *info = lay.info;
    return rc;
}

int eddy_iovector_read(const char* path, double* u, double* v, double* w, int nx, int ny, int nz, int myrank,
                       int mysize, eddy_field_info* info) {
    return iovector_read_impl(path, u, v, w, nx, ny, nz, myrank, mysize, info, false, nullptr);
}

int eddy_iovector_read_dev(const char* path, double* d_u, double* d_v, double* d_w, int nx, int ny, int nz,
                           int myrank, int mysize, eddy_field_info* info, void* stream) {
    return iovector_read_impl(path, d_u, d_v, d_w, nx, ny, nz, myrank, mysize, info, true, stream);
}

static int ioscalar_write_impl(const char* path, const double* p, int nx, int ny, int nz, int per_rank, int nstep,
                               double time, double dtm1, double grav, bool device, void* stream) {
    const int rc = check_write_args(path, p, nx, ny, nz);
    if (rc) return rc;
    Header hd{{nx, ny, nz, 1}, {time, dtm1, grav}, per_rank ? 3 : 0};
    const int32_t ns = nstep;
    const double dg[2] = {dtm1, grav};
    const double* src[1] = {p};
    return write_field(path, hd, src, 1, nz, device, (hipStream_t)stream, per_rank ? nullptr : &ns, &time, dg);
}

int eddy_ioscalar_write(const char* path, const double* p, int nx, int ny, int nz, int per_rank, int nstep,
                        double time, double dtm1, double grav) {
    return ioscalar_write_impl(path, p, nx, ny, nz, per_rank, nstep, time, dtm1, grav, false, nullptr);
}

int eddy_ioscalar_write_dev(const char* path, const double* d_p, int nx, int ny, int nz, int per_rank, int nstep,
                            double time, double dtm1, double grav, void* stream) {
    return ioscalar_write_impl(path, d_p, nx, ny, nz, per_rank, nstep, time, dtm1, grav, true, stream);
}

static int iovector_write_impl(const char* path, const double* u, const double* v, const double* w, int nx, int ny,
                               int nz, int per_rank, double time, bool device, void* stream) {
    int rc = check_write_args(path, u, nx, ny, nz);
    if (!rc) rc = check_write_args(path, v, nx, ny, nz);
    if (!rc) rc = check_write_args(path, w, nx, ny, nz);
    if (rc) return rc;
    Header hd{{nx, ny, nz, 3}, {time, 0.0, 0.0}, per_rank ? 1 : 0};
    const double* src[3] = {u, v, w};
    return write_field(path, hd, src, 3, nz, device, (hipStream_t)stream, nullptr, nullptr, nullptr);
}

int eddy_iovector_write(const char* path, const double* u, const double* v, const double* w, int nx, int ny, int nz,
                        int per_rank, double time) {
    return iovector_write_impl(path, u, v, w, nx, ny, nz, per_rank, time, false, nullptr);
}

int eddy_iovector_write_dev(const char* path, const double* d_u, const double* d_v, const double* d_w, int nx,
                            int ny, int nz, int per_rank, double time, void* stream) {
    return iovector_write_impl(path, d_u, d_v, d_w, nx, ny, nz, per_rank, time, true, stream);
}

}  // extern "C"
