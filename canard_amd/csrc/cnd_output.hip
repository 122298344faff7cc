// Output of results: the Tecplot binary writer of write_output_file /
// techead (src/mo_io.f90:345-529), the output variables and schedule of the
// time loop (src/mo_canard_driver.f90:173-195, 286-291, 373-425), and
// asynchronous writing: the output variables are formed on the device, copied
// to pinned host memory on a second stream and written by a host thread while
// the context's stream goes on stepping -- the counterpart of the reference's
// I/O server ranks (src/mo_io_server.f90:224-261).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "cnd_ctx.hpp"

// Buffers and the writer thread of a context's output (created on first use).
struct OutputState {
    float *dev = nullptr;    // output variables on the device
    float *host = nullptr;   // pinned copy the writer thread reads
    size_t cap = 0;          // floats in each
    hipStream_t st = nullptr;                          // copy stream
    hipEvent_t formed = nullptr, copied = nullptr;
    std::thread writer;
    bool busy = false;
    int rc = CND_OK;
    std::string err;
};

namespace {

// Output variables of one block (mo_canard_driver.f90:379-400): the time
// average qb / dtsum (the filtered state qo after the first step, n = 1)
// as density, velocity in the fixed frame and pressure, in FP32, stored
// variable after variable over the block's points.
__global__ __launch_bounds__(kPts) void k_record_out(const BlkDev *__restrict__ blks,
                                                    const InstDev *__restrict__ inst, int b,
                                                    const double *__restrict__ qo, const double *__restrict__ qb,
                                                    float *__restrict__ out, int64_t NT)
{
    const BlkDev &B = blks[b];
    const InstDev &I = inst[B.inst];
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t lp = int64_t(blockIdx.x) * kPts + threadIdx.x;
    if (lp >= npt) return;
    const int64_t g = B.off + lp;
    double q[5];
    if (I.n == 1) {
#pragma unroll
        for (int m = 0; m < 5; ++m) q[m] = qo[m * NT + g];
    } else {
        const double ra0 = 1.0 / I.dtsum;
#pragma unroll
        for (int m = 0; m < 5; ++m) q[m] = ra0 * qb[m * NT + g];
    }
    const double rr = 1.0 / q[0];
    out[lp] = float(q[0]);
#pragma unroll
    for (int m = 1; m < 4; ++m) out[m * npt + lp] = float(rr * q[m] + I.umf[m - 1]);
    out[4 * npt + lp] = float(kGm1 * (q[4] - (0.5 * rr) * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3])));
}

// Grid of one block in FP32, x, y, z after each other (mo_canard_driver.f90:180-186).
__global__ __launch_bounds__(kPts) void k_grid_out(const BlkDev *__restrict__ blks, int b,
                                                  const double *__restrict__ xyz, float *__restrict__ out,
                                                  int64_t NT)
{
    const BlkDev &B = blks[b];
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t lp = int64_t(blockIdx.x) * kPts + threadIdx.x;
    if (lp >= npt) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c * npt + lp] = float(xyz[c * NT + B.off + lp]);
}

// The reset after an output (mo_canard_driver.f90:419-420): qb = 0 on the
// points of instance inst's blocks, and its dtsum.
__global__ __launch_bounds__(kPts) void k_record_clear(const BlkDev *__restrict__ blks, InstDev *__restrict__ inst,
                                                      const Tile *__restrict__ tiles, int inst_id,
                                                      double *__restrict__ qb, int64_t NT)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) inst[inst_id].dtsum = 0.0;
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    if (B.inst != inst_id) return;
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t lp = int64_t(t.q0) + threadIdx.x;
    if (lp >= npt) return;
#pragma unroll
    for (int m = 0; m < 5; ++m) qb[m * NT + B.off + lp] = 0.0;
}

// ---- Tecplot binary (#!TDV112) ------------------------------------------------
// The layout of techead (mo_io.f90:417-513) followed by each zone's FP32 data:
// file header (magic, byte order, file type, title, variable names), one
// header record per zone (marker 299.0, name, parent -1, strand, solution
// time, -1, ordered zone type 0, node location 0, no face neighbours, I, J,
// K maxima, no auxiliary data), end-of-header marker 357.0; then per zone a
// data record (marker 299.0, FP32 format per variable, no passive or shared
// variables, share zone -1, FP64 minimum and maximum of every variable) and
// the values, variable after variable, i fastest.  Strings are one 32-bit
// integer per character and a terminating 0.
struct Bytes {
    std::vector<unsigned char> b;
    void raw(const void *p, size_t n)
    {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        b.insert(b.end(), c, c + n);
    }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
    void f64(double v) { raw(&v, 8); }
    void str(const std::string &s)
    {
        for (unsigned char ch : s) i32(int32_t(ch));
        i32(0);
    }
};

std::string zone_name(int z)
{
    char buf[16];
    std::snprintf(buf, sizeof buf, "z%03d", z % 1000);
    return buf;
}

// Writes the file; returns an empty string or the error.
std::string write_plt(const std::string &path, const std::string &title, int file_type,
                      const std::vector<std::string> &names, const std::vector<const int *> &dims,
                      double time, const std::vector<const float *> &data)
{
    const int nvar = int(names.size()), nz = int(dims.size());
    Bytes h;
    h.raw("#!TDV112", 8);
    h.i32(1);
    h.i32(file_type);
    h.str(title);
    h.i32(nvar);
    for (const std::string &s : names) h.str(s);
    for (int z = 0; z < nz; ++z) {
        h.f32(299.0f);
        h.str(zone_name(z));
        h.i32(-1);
        h.i32(z + 1);
        h.f64(time);
        h.i32(-1);
        h.i32(0);
        h.i32(0);
        h.i32(0);
        h.i32(0);
        for (int k = 0; k < 3; ++k) h.i32(dims[z][k]);
        h.i32(0);
    }
    h.f32(357.0f);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return "cannot open " + path + " for writing";
    bool ok = std::fwrite(h.b.data(), 1, h.b.size(), f) == h.b.size();
    for (int z = 0; ok && z < nz; ++z) {
        const size_t npt = size_t(dims[z][0]) * dims[z][1] * dims[z][2];
        Bytes r;
        r.f32(299.0f);
        for (int m = 0; m < nvar; ++m) r.i32(1);
        r.i32(0);
        r.i32(0);
        r.i32(-1);
        for (int m = 0; m < nvar; ++m) {   // vminmax (mo_io.f90:533-562) over the FP32 values
            const float *v = data[z] + m * npt;
            float lo = v[0], hi = v[0];
            for (size_t i = 1; i < npt; ++i) { lo = std::min(lo, v[i]); hi = std::max(hi, v[i]); }
            r.f64(double(lo));
            r.f64(double(hi));
        }
        ok = std::fwrite(r.b.data(), 1, r.b.size(), f) == r.b.size() &&
             std::fwrite(data[z], sizeof(float), nvar * npt, f) == nvar * npt;
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? std::string() : "write to " + path + " failed";
}

OutputState *out_state(cnd_ctx *c)
{
    if (!c->out) c->out = new OutputState();
    return c->out;
}

// Room for n floats in the device and pinned host buffers (after any write
// in flight has finished with them).
int out_room(cnd_ctx *c, size_t n)
{
    OutputState *o = out_state(c);
    if (!o->st) {
        HIPX(c, hipStreamCreateWithFlags(&o->st, hipStreamNonBlocking));
        HIPX(c, hipEventCreateWithFlags(&o->formed, hipEventDisableTiming));
        HIPX(c, hipEventCreateWithFlags(&o->copied, hipEventDisableTiming));
    }
    if (o->cap >= n) return CND_OK;
    if (o->dev) { HIPX(c, hipFree(o->dev)); o->dev = nullptr; }
    if (o->host) { HIPX(c, hipHostFree(o->host)); o->host = nullptr; }
    o->cap = 0;
    HIPX(c, hipMalloc((void **)&o->dev, n * sizeof(float)));
    HIPX(c, hipHostMalloc((void **)&o->host, n * sizeof(float), hipHostMallocDefault));
    o->cap = n;
    return CND_OK;
}

// Wait for the writer thread; its status.
int out_join(cnd_ctx *c)
{
    OutputState *o = c->out;
    if (!o || !o->busy) return CND_OK;
    o->writer.join();
    o->busy = false;
    if (o->rc != CND_OK) {
        const int rc = o->rc;
        o->rc = CND_OK;
        return fail(c, rc, o->err);
    }
    return CND_OK;
}

int64_t block_points(const cnd_ctx *c, int b)
{
    return int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
}

}  // namespace

void cnd::output_release(cnd_ctx *c)
{
    OutputState *o = c->out;
    if (!o) return;
    if (o->busy) o->writer.join();
    if (o->st) (void)hipStreamSynchronize(o->st);
    if (o->dev) (void)hipFree(o->dev);
    if (o->host) (void)hipHostFree(o->host);
    if (o->formed) (void)hipEventDestroy(o->formed);
    if (o->copied) (void)hipEventDestroy(o->copied);
    if (o->st) (void)hipStreamDestroy(o->st);
    delete o;
    c->out = nullptr;
}

extern "C" {

int cnd_write_plt(const char *path, const char *title, int file_type, int nvar, const char *const *names,
                  int nzones, const int *dims, double time, const float *const *data)
{
    if (!path || !title || !names || !dims || !data || nvar < 1 || nzones < 1 || (file_type != 1 && file_type != 2))
        return fail(nullptr, CND_EINVAL, "cnd_write_plt: bad argument");
    std::vector<std::string> nm;
    for (int m = 0; m < nvar; ++m) {
        if (!names[m]) return fail(nullptr, CND_EINVAL, "cnd_write_plt: null variable name");
        nm.emplace_back(names[m]);
    }
    std::vector<const int *> dz;
    std::vector<const float *> pz;
    for (int z = 0; z < nzones; ++z) {
        if (!data[z] || dims[3 * z] < 1 || dims[3 * z + 1] < 1 || dims[3 * z + 2] < 1)
            return fail(nullptr, CND_EINVAL, "cnd_write_plt: bad zone " + std::to_string(z));
        dz.push_back(dims + 3 * z);
        pz.push_back(data[z]);
    }
    const std::string e = write_plt(path, title, file_type, nm, dz, time, pz);
    return e.empty() ? CND_OK : fail(nullptr, CND_EDEVICE, "cnd_write_plt: " + e);
}

int cnd_output_schedule(cnd_ctx *c, int inst, double tmax, int ndata)
{
    if (!c || inst < 0 || inst >= c->ninst || (ndata > 0 && !(tmax > 0.0)))
        return fail(c, CND_EINVAL, "cnd_output_schedule: bad argument");
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_output_schedule: a step is half done");
    HIPX(c, hipStreamSynchronize(c->st));
    InstDev I;
    HIPX(c, hipMemcpy(&I, c->dinst + inst, sizeof I, hipMemcpyDeviceToHost));
    I.tmax = tmax;
    I.ndata = ndata > 0 ? ndata : 0;
    I.ndati = -1;
    I.nout = 0;
    HIPX(c, hipMemcpy(c->dinst + inst, &I, sizeof I, hipMemcpyHostToDevice));
    return CND_OK;
}

int cnd_output_due(cnd_ctx *c, int inst, int *due, int *ndati, double *time)
{
    if (!c || inst < 0 || inst >= c->ninst) return fail(c, CND_EINVAL, "cnd_output_due: bad instance");
    HIPX(c, hipStreamSynchronize(c->st));
    InstDev I;
    HIPX(c, hipMemcpy(&I, c->dinst + inst, sizeof I, hipMemcpyDeviceToHost));
    if (due) *due = I.nout;
    if (ndati) *ndati = I.ndati;
    if (time) *time = I.timo - 0.5 * I.dtsum;   // times(ndati) = timo - dtsum / 2
    return CND_OK;
}

int cnd_output_write(cnd_ctx *c, int inst, const char *path, int async)
{
    if (!c || !path || inst < 0 || inst >= c->ninst) return fail(c, CND_EINVAL, "cnd_output_write: bad argument");
    if (!c->qb) return fail(c, CND_ESTATE, "cnd_output_write: recording was never enabled (cnd_record)");
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_output_write: a step is half done");
    int r;
    if ((r = out_join(c))) return r;
    // the output's number and time, from the instance's clock after the step
    HIPX(c, hipStreamSynchronize(c->st));
    InstDev I;
    HIPX(c, hipMemcpy(&I, c->dinst + inst, sizeof I, hipMemcpyDeviceToHost));
    std::vector<int> blocks;
    size_t total = 0;
    for (int b = 0; b < c->nblk; ++b)
        if (c->spec[b].inst == inst) { blocks.push_back(b); total += 5 * size_t(block_points(c, b)); }
    if ((r = out_room(c, total))) return r;
    OutputState *o = c->out;
    size_t at = 0;
    std::vector<size_t> offs;
    for (int b : blocks) {
        const int64_t n = block_points(c, b);
        hipLaunchKernelGGL(k_record_out, dim3(unsigned((n + kPts - 1) / kPts)), dim3(kPts), 0, c->st, c->dblk,
                           c->dinst, b, c->qo, c->qb, o->dev + at, c->NT);
        offs.push_back(at);
        at += 5 * size_t(n);
    }
    hipLaunchKernelGGL(k_record_clear, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt, inst, c->qb,
                       c->NT);
    if ((r = check_launch(c))) return r;
    HIPX(c, hipEventRecord(o->formed, c->st));
    HIPX(c, hipStreamWaitEvent(o->st, o->formed, 0));
    HIPX(c, hipMemcpyAsync(o->host, o->dev, total * sizeof(float), hipMemcpyDeviceToHost, o->st));
    HIPX(c, hipEventRecord(o->copied, o->st));
    char title[16];
    std::snprintf(title, sizeof title, "n%03d", ((I.ndati % 1000) + 1000) % 1000);
    const double time = I.timo - 0.5 * I.dtsum;
    std::vector<int> dims;
    for (int b : blocks)
        for (int d = 0; d < 3; ++d) dims.push_back(c->spec[b].n[d]);
    auto job = [o, path = std::string(path), title = std::string(title), dims, offs, time]() {
        if (hipEventSynchronize(o->copied) != hipSuccess) {
            o->rc = CND_EDEVICE;
            o->err = "cnd_output_write: device to host copy failed";
            return;
        }
        std::vector<const int *> dz;
        std::vector<const float *> pz;
        for (size_t z = 0; z < offs.size(); ++z) {
            dz.push_back(dims.data() + 3 * z);
            pz.push_back(o->host + offs[z]);
        }
        const std::string e = write_plt(path, title, 2, {"rho", "u", "v", "w", "p"}, dz, time, pz);
        if (!e.empty()) { o->rc = CND_EDEVICE; o->err = "cnd_output_write: " + e; }
    };
    if (async) {
        o->writer = std::thread(job);
        o->busy = true;
        return CND_OK;
    }
    job();
    if (o->rc != CND_OK) { o->rc = CND_OK; return fail(c, CND_EDEVICE, o->err); }
    return CND_OK;
}

int cnd_output_grid(cnd_ctx *c, int inst, const char *path)
{
    if (!c || !path || inst < 0 || inst >= c->ninst) return fail(c, CND_EINVAL, "cnd_output_grid: bad argument");
    for (int b = 0; b < c->nblk; ++b)
        if (c->spec[b].inst == inst && !c->grid_set[b])
            return fail(c, CND_ESTATE, "cnd_output_grid: grid of block " + std::to_string(b) + " not set");
    int r;
    if ((r = out_join(c))) return r;
    std::vector<int> blocks, dims;
    std::vector<size_t> offs;
    size_t total = 0;
    for (int b = 0; b < c->nblk; ++b)
        if (c->spec[b].inst == inst) {
            blocks.push_back(b);
            offs.push_back(total);
            total += 3 * size_t(block_points(c, b));
            for (int d = 0; d < 3; ++d) dims.push_back(c->spec[b].n[d]);
        }
    if ((r = out_room(c, total))) return r;
    OutputState *o = c->out;
    for (size_t z = 0; z < blocks.size(); ++z) {
        const int64_t n = block_points(c, blocks[z]);
        hipLaunchKernelGGL(k_grid_out, dim3(unsigned((n + kPts - 1) / kPts)), dim3(kPts), 0, c->st, c->dblk, blocks[z],
                           c->xyz, o->dev + offs[z], c->NT);
    }
    if ((r = check_launch(c))) return r;
    HIPX(c, hipMemcpyAsync(o->host, o->dev, total * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    std::vector<const int *> dz;
    std::vector<const float *> pz;
    for (size_t z = 0; z < blocks.size(); ++z) {
        dz.push_back(dims.data() + 3 * z);
        pz.push_back(o->host + offs[z]);
    }
    const std::string e = write_plt(path, "grid", 1, {"x", "y", "z"}, dz, 0.0, pz);
    return e.empty() ? CND_OK : fail(c, CND_EDEVICE, "cnd_output_grid: " + e);
}

int cnd_output_wait(cnd_ctx *c)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    return out_join(c);
}

int cnd_record_read(cnd_ctx *c, int b, float *out)
{
    if (!c || !out || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_record_read: bad argument");
    if (!c->qb) return fail(c, CND_ESTATE, "cnd_record_read: recording was never enabled (cnd_record)");
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_record_read: a step is half done");
    int r;
    if ((r = out_join(c))) return r;
    const int64_t n = block_points(c, b);
    if ((r = out_room(c, 5 * size_t(n)))) return r;
    hipLaunchKernelGGL(k_record_out, dim3(unsigned((n + kPts - 1) / kPts)), dim3(kPts), 0, c->st, c->dblk, c->dinst, b,
                       c->qo, c->qb, c->out->dev, c->NT);
    if ((r = check_launch(c))) return r;
    HIPX(c, hipMemcpyAsync(out, c->out->dev, sizeof(float) * 5 * n, hipMemcpyDeviceToHost, c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

}  // extern "C"
