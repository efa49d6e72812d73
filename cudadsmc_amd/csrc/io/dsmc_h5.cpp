// dsmc_h5.cpp -- HDF5 output in the layout of the reference's writer
// (CUDADSMC/src/hdf5Helpers.cu:13-163, datasets of main.cu:202-383 and
// main.cu:526-576).  Host code; links libhdf5 (1.10 C API).
//
// The reference opens and closes the file for every dataset write; this
// writer keeps the file and its datasets open and flushes after each frame,
// which leaves the same file on disk.
#include "../../../include/dsmc_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

dsmc_status fail(const std::string &msg, dsmc_status st = DSMC_EINVAL)
{
    g_err = msg;
    return st;
}

hid_t native(int type)
{
    switch (type) {
    case DSMC_H5_DOUBLE: return H5T_NATIVE_DOUBLE;
    case DSMC_H5_INT: return H5T_NATIVE_INT;
    case DSMC_H5_BOOL: return H5T_NATIVE_HBOOL;
    default: return -1;
    }
}

size_t elem_bytes(int type)
{
    return type == DSMC_H5_DOUBLE ? sizeof(double) : type == DSMC_H5_INT ? sizeof(int)
                                                                          : sizeof(hbool_t);
}

struct Dataset {
    hid_t id = -1;
    hsize_t rows = 0, cols = 0, frames = 0;
    int type = 0;
};

}  // namespace

struct dsmc_h5 {
    hid_t file = -1;
    std::string group;
    std::map<std::string, Dataset> sets;
    bool frame_sets = false;   // datasets of dsmc_h5_write_frame created
};

extern "C" {

const char *dsmc_h5_last_error(void) { return g_err.c_str(); }

dsmc_status dsmc_h5_create(const char *filename, const char *group, dsmc_h5 **out)
{
    if (!filename || !group || !out) return fail("dsmc_h5_create: null argument");
    *out = nullptr;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are reported by status
    const hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot create ") + filename);
    const hid_t g = H5Gcreate2(f, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) {
        H5Fclose(f);
        return fail(std::string("cannot create group ") + group);
    }
    H5Gclose(g);
    dsmc_h5 *h = new dsmc_h5;
    h->file = f;
    h->group = group;
    *out = h;
    return DSMC_OK;
}

dsmc_status dsmc_h5_dataset(dsmc_h5 *h, const char *name, long long rows, long long cols,
                            int type)
{
    if (!h || !name) return fail("dsmc_h5_dataset: null argument");
    if (rows < 1 || cols < 1) return fail("dsmc_h5_dataset: rows and cols must be >= 1");
    if (native(type) < 0) return fail("dsmc_h5_dataset: unknown type");
    if (h->sets.count(name)) return fail(std::string("dataset exists: ") + name);
    // (rows, cols, t): t starts at 0 and is unlimited; one chunk per frame
    const hsize_t dims[3] = { (hsize_t)rows, (hsize_t)cols, 0 };
    const hsize_t maxdims[3] = { (hsize_t)rows, (hsize_t)cols, H5S_UNLIMITED };
    const hsize_t chunk[3] = { (hsize_t)rows, (hsize_t)cols, 1 };
    const hid_t space = H5Screate_simple(3, dims, maxdims);
    const hid_t prop = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(prop, 3, chunk);
    const hid_t ds = H5Dcreate2(h->file, name, native(type), space, H5P_DEFAULT, prop, H5P_DEFAULT);
    H5Pclose(prop);
    H5Sclose(space);
    if (ds < 0) return fail(std::string("cannot create dataset ") + name);
    Dataset d;
    d.id = ds;
    d.rows = (hsize_t)rows;
    d.cols = (hsize_t)cols;
    d.type = type;
    h->sets[name] = d;
    return DSMC_OK;
}

dsmc_status dsmc_h5_append(dsmc_h5 *h, const char *name, const void *data)
{
    if (!h || !name || !data) return fail("dsmc_h5_append: null argument");
    auto it = h->sets.find(name);
    if (it == h->sets.end()) return fail(std::string("no dataset ") + name);
    Dataset &d = it->second;
    const hsize_t ext[3] = { d.rows, d.cols, d.frames + 1 };
    if (H5Dset_extent(d.id, ext) < 0) return fail(std::string("cannot extend ") + name);
    const hid_t fspace = H5Dget_space(d.id);
    const hsize_t off[3] = { 0, 0, d.frames };
    const hsize_t cnt[3] = { d.rows, d.cols, 1 };
    H5Sselect_hyperslab(fspace, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    const hid_t mspace = H5Screate_simple(3, cnt, nullptr);
    const herr_t st = H5Dwrite(d.id, native(d.type), mspace, fspace, H5P_DEFAULT, data);
    H5Sclose(mspace);
    H5Sclose(fspace);
    if (st < 0) return fail(std::string("cannot write ") + name);
    ++d.frames;
    H5Fflush(h->file, H5F_SCOPE_LOCAL);
    return DSMC_OK;
}

void dsmc_h5_close(dsmc_h5 *h)
{
    if (!h) return;
    for (auto &kv : h->sets) H5Dclose(kv.second.id);
    if (h->file >= 0) {
        H5Fflush(h->file, H5F_SCOPE_LOCAL);
        H5Fclose(h->file);
    }
    delete h;
}

dsmc_status dsmc_h5_write_frame(dsmc_h5 *h, dsmc_ctx *ctx, double time)
{
    if (!h || !ctx) return fail("dsmc_h5_write_frame: null argument");
    dsmc_info info;
    dsmc_status st = dsmc_get_info(ctx, &info);
    if (st != DSMC_OK) return fail(std::string("dsmc_get_info: ") + dsmc_last_error(ctx), st);
    const size_t N = (size_t)info.n_capacity, n = (size_t)info.n_alive;
    const std::string g = h->group;
    // per-cell arrays: cells + 1 entries in the FP64 form (main.cu:300-331)
    size_t cell_bytes = 0;
    const bool have_cells = dsmc_field_bytes(ctx, DSMC_FIELD_CELL_COUNT, &cell_bytes) == DSMC_OK &&
                            cell_bytes > 0;
    const size_t C1 = have_cells ? cell_bytes / sizeof(int) : 1;
    if (!h->frame_sets) {
        const struct { const char *name; long long rows, cols; int type; } sets[] = {
            { "positions", (long long)N, 3, DSMC_H5_DOUBLE },
            { "velocities", (long long)N, 3, DSMC_H5_DOUBLE },
            { "evapPos", (long long)N, 3, DSMC_H5_DOUBLE },
            { "evapVel", (long long)N, 3, DSMC_H5_DOUBLE },
            { "psiUp", (long long)N, 2, DSMC_H5_DOUBLE },
            { "psiDn", (long long)N, 2, DSMC_H5_DOUBLE },
            { "collisionCount", (long long)C1, 1, DSMC_H5_INT },
            { "atomCount", (long long)C1, 1, DSMC_H5_INT },
            { "simuatedTime", 1, 1, DSMC_H5_DOUBLE },   // the reference's spelling
            { "atomNumber", 1, 1, DSMC_H5_INT },
            { "cellID", (long long)N, 1, DSMC_H5_INT },
            { "atomID", (long long)N, 1, DSMC_H5_INT },
            { "atomIsSpinUp", (long long)N, 1, DSMC_H5_BOOL },
        };
        for (const auto &s : sets) {
            st = dsmc_h5_dataset(h, (g + "/" + s.name).c_str(), s.rows, s.cols, s.type);
            if (st != DSMC_OK) return st;
        }
        h->frame_sets = true;
    }
    auto copy = [&](int field, void *dst, size_t bytes) -> dsmc_status {
        if (bytes == 0) return DSMC_OK;   // no alive atoms
        const dsmc_status s = dsmc_copy_out(ctx, field, dst, bytes);
        if (s != DSMC_OK) g_err = std::string("dsmc_copy_out: ") + dsmc_last_error(ctx);
        return s;
    };
    auto put = [&](const char *name, const void *data) {
        return dsmc_h5_append(h, (g + "/" + name).c_str(), data);
    };
    std::vector<double> d3(3 * N, 0.0);
    // positions, velocities: alive rows, zeros after
    if ((st = copy(DSMC_FIELD_POS, d3.data(), 3 * n * sizeof(double))) != DSMC_OK) return st;
    std::fill(d3.begin() + 3 * n, d3.end(), 0.0);
    if ((st = put("positions", d3.data())) != DSMC_OK) return st;
    if ((st = copy(DSMC_FIELD_VEL, d3.data(), 3 * n * sizeof(double))) != DSMC_OK) return st;
    if ((st = put("velocities", d3.data())) != DSMC_OK) return st;
    // removed atoms, indexed by label
    if ((st = copy(DSMC_FIELD_EVAP_POS, d3.data(), 3 * N * sizeof(double))) != DSMC_OK) return st;
    if ((st = put("evapPos", d3.data())) != DSMC_OK) return st;
    if ((st = copy(DSMC_FIELD_EVAP_VEL, d3.data(), 3 * N * sizeof(double))) != DSMC_OK) return st;
    if ((st = put("evapVel", d3.data())) != DSMC_OK) return st;
    // spinors (re, im) of spin up and down
    std::vector<double> up(2 * N, 0.0), dn(2 * N, 0.0);
    size_t spin_bytes = 0;
    if (dsmc_field_bytes(ctx, DSMC_FIELD_SPIN, &spin_bytes) == DSMC_OK && spin_bytes > 0 && n > 0) {
        const size_t comp = spin_bytes / (n * sizeof(double));   // 4 (spinor) or 3 (Bloch)
        std::vector<double> s(comp * n);
        if ((st = copy(DSMC_FIELD_SPIN, s.data(), s.size() * sizeof(double))) != DSMC_OK) return st;
        for (size_t i = 0; i < n; ++i) {
            const double *a = s.data() + comp * i;
            if (comp == 4) {
                up[2 * i] = a[0]; up[2 * i + 1] = a[1];
                dn[2 * i] = a[2]; dn[2 * i + 1] = a[3];
            } else {
                // Bloch vector (x, y, z) -> (cos(th/2), e^{i phi} sin(th/2))
                const double th = std::acos(std::fmax(-1.0, std::fmin(1.0, a[2])));
                const double ph = std::atan2(a[1], a[0]);
                up[2 * i] = std::cos(0.5 * th);
                dn[2 * i] = std::sin(0.5 * th) * std::cos(ph);
                dn[2 * i + 1] = std::sin(0.5 * th) * std::sin(ph);
            }
        }
    }
    if ((st = put("psiUp", up.data())) != DSMC_OK) return st;
    if ((st = put("psiDn", dn.data())) != DSMC_OK) return st;
    // per-cell collision and atom counts
    std::vector<int> cells(C1, 0);
    if (have_cells) {
        size_t cb = 0;
        if (dsmc_field_bytes(ctx, DSMC_FIELD_COLLISION_COUNT, &cb) == DSMC_OK && cb == C1 * sizeof(int))
            if ((st = copy(DSMC_FIELD_COLLISION_COUNT, cells.data(), cb)) != DSMC_OK) return st;
    }
    if ((st = put("collisionCount", cells.data())) != DSMC_OK) return st;
    std::fill(cells.begin(), cells.end(), 0);
    if (have_cells)
        if ((st = copy(DSMC_FIELD_CELL_COUNT, cells.data(), C1 * sizeof(int))) != DSMC_OK) return st;
    if ((st = put("atomCount", cells.data())) != DSMC_OK) return st;
    if ((st = put("simuatedTime", &time)) != DSMC_OK) return st;
    const int alive = (int)n;
    if ((st = put("atomNumber", &alive)) != DSMC_OK) return st;
    std::vector<int> iv(N, 0);
    if ((st = copy(DSMC_FIELD_CELL_ID, iv.data(), n * sizeof(int))) != DSMC_OK) return st;
    if ((st = put("cellID", iv.data())) != DSMC_OK) return st;
    std::fill(iv.begin(), iv.end(), 0);
    if ((st = copy(DSMC_FIELD_ATOM_ID, iv.data(), n * sizeof(int))) != DSMC_OK) return st;
    if ((st = put("atomID", iv.data())) != DSMC_OK) return st;
    // spin state of the alive atoms (flipped atoms are down in the spin-flip model)
    std::fill(iv.begin(), iv.end(), 0);
    if ((st = copy(DSMC_FIELD_SPIN_UP, iv.data(), n * sizeof(int))) != DSMC_OK) return st;
    static_assert(sizeof(hbool_t) == 1, "one byte per flag");
    std::vector<unsigned char> sp(N, 0);
    for (size_t i = 0; i < n; ++i) sp[i] = iv[i] ? 1 : 0;
    return put("atomIsSpinUp", sp.data());
}

dsmc_status dsmc_h5_dims(const char *filename, const char *name, long long dims[3])
{
    if (!filename || !name || !dims) return fail("dsmc_h5_dims: null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot open ") + filename);
    const hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return fail(std::string("no dataset ") + name);
    }
    const hid_t sp = H5Dget_space(d);
    hsize_t dd[3] = { 0, 0, 0 };
    const int rank = H5Sget_simple_extent_ndims(sp);
    dsmc_status st = DSMC_OK;
    if (rank != 3) st = fail(std::string("dataset is not rank 3: ") + name);
    else H5Sget_simple_extent_dims(sp, dd, nullptr);
    for (int k = 0; k < 3; ++k) dims[k] = (long long)dd[k];
    H5Sclose(sp);
    H5Dclose(d);
    H5Fclose(f);
    return st;
}

dsmc_status dsmc_h5_read(const char *filename, const char *name, int type, void *out,
                         size_t bytes)
{
    long long dims[3];
    dsmc_status st = dsmc_h5_dims(filename, name, dims);
    if (st != DSMC_OK) return st;
    if (native(type) < 0) return fail("dsmc_h5_read: unknown type");
    const size_t need = (size_t)(dims[0] * dims[1] * dims[2]) * elem_bytes(type);
    if (!out || bytes < need) return fail("dsmc_h5_read: buffer too small");
    const hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    const hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    const herr_t r = H5Dread(d, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(d);
    H5Fclose(f);
    return r < 0 ? fail(std::string("cannot read ") + name) : DSMC_OK;
}

}  // extern "C"
