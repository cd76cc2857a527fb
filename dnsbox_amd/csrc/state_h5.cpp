// State files of the reference: stateNNNN.h5 with datasets u, v, w, the physical velocity in
// FP64, Fortran shape (Nx, Ny, Nz) (io_save_state, io.f90:100-278; io_load_state,
// io.f90:455-626).  A Fortran dataset of shape (Nx, Ny, Nz) is the C-order dataspace
// {Nz, Ny, Nx} with x fastest, which is the byte order of u(Nx, Ny, Nz) and of the physical
// arrays of dnsbox_get_state_physical / dnsbox_set_state_physical.  The reference writes each
// dataset chunked by its ranks' slabs (io.f90:52-55); one writer here chunks along z in slabs of
// at most 64 MiB.  Host code over the HDF5 C library; no GPU.
#include <hdf5.h>

#include <string>

#include "../../include/dnsbox_io.h"

namespace {

thread_local std::string g_h5err;

int fail(int code, const std::string& msg)
{
    g_h5err = msg;
    return code;
}

const char* const kNames[3] = {"u", "v", "w"};

// closes an HDF5 handle when it leaves scope
struct H5Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Handle()
    {
        if (id >= 0) close(id);
    }
    H5Handle(const H5Handle&) = delete;
    H5Handle& operator=(const H5Handle&) = delete;
};

bool bad_dims(int nx, int ny, int nz) { return nx < 1 || ny < 1 || nz < 1; }

}  // namespace

extern "C" {

const char* dnsbox_h5_last_error(void) { return g_h5err.c_str(); }

int dnsbox_h5_write_state(const char* path, int nx, int ny, int nz, const double* u, const double* v,
                          const double* w)
{
    if (!path || !u || !v || !w) return fail(DNSBOX_IO_E_ARG, "null argument");
    if (bad_dims(nx, ny, nz)) return fail(DNSBOX_IO_E_ARG, "grid dimensions must be >= 1");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are returned, not printed
    H5Handle file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(DNSBOX_IO_E_IO, std::string("cannot create ") + path);
    const hsize_t dims[3] = {hsize_t(nz), hsize_t(ny), hsize_t(nx)};
    const hsize_t plane = hsize_t(nx) * hsize_t(ny) * sizeof(double);
    hsize_t cz = (hsize_t(64) << 20) / plane;
    if (cz < 1) cz = 1;
    if (cz > hsize_t(nz)) cz = hsize_t(nz);
    const hsize_t chunk[3] = {cz, hsize_t(ny), hsize_t(nx)};
    H5Handle space(H5Screate_simple(3, dims, nullptr), H5Sclose);
    H5Handle dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (space.id < 0 || dcpl.id < 0 || H5Pset_chunk(dcpl.id, 3, chunk) < 0)
        return fail(DNSBOX_IO_E_IO, "cannot set up the dataspace");
    const double* src[3] = {u, v, w};
    for (int f = 0; f < 3; ++f) {
        H5Handle ds(H5Dcreate2(file.id, kNames[f], H5T_IEEE_F64LE, space.id, H5P_DEFAULT, dcpl.id, H5P_DEFAULT),
                    H5Dclose);
        if (ds.id < 0) return fail(DNSBOX_IO_E_IO, std::string("cannot create dataset ") + kNames[f]);
        if (H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, src[f]) < 0)
            return fail(DNSBOX_IO_E_IO, std::string("cannot write dataset ") + kNames[f]);
    }
    if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return fail(DNSBOX_IO_E_IO, "flush failed");
    return 0;
}

int dnsbox_h5_state_dims(const char* path, int* nx, int* ny, int* nz)
{
    if (!path || !nx || !ny || !nz) return fail(DNSBOX_IO_E_ARG, "null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Handle file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(DNSBOX_IO_E_IO, std::string("cannot open ") + path);
    H5Handle ds(H5Dopen2(file.id, "u", H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) return fail(DNSBOX_IO_E_IO, "no dataset u");
    H5Handle space(H5Dget_space(ds.id), H5Sclose);
    hsize_t d[3] = {0, 0, 0};
    if (H5Sget_simple_extent_ndims(space.id) != 3 || H5Sget_simple_extent_dims(space.id, d, nullptr) != 3)
        return fail(DNSBOX_IO_E_IO, "dataset u is not three-dimensional");
    *nz = int(d[0]);
    *ny = int(d[1]);
    *nx = int(d[2]);
    return 0;
}

int dnsbox_h5_read_state(const char* path, int nx, int ny, int nz, double* u, double* v, double* w)
{
    if (!path || !u || !v || !w) return fail(DNSBOX_IO_E_ARG, "null argument");
    if (bad_dims(nx, ny, nz)) return fail(DNSBOX_IO_E_ARG, "grid dimensions must be >= 1");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Handle file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(DNSBOX_IO_E_IO, std::string("cannot open ") + path);
    double* dst[3] = {u, v, w};
    for (int f = 0; f < 3; ++f) {
        H5Handle ds(H5Dopen2(file.id, kNames[f], H5P_DEFAULT), H5Dclose);
        if (ds.id < 0) return fail(DNSBOX_IO_E_IO, std::string("no dataset ") + kNames[f]);
        H5Handle space(H5Dget_space(ds.id), H5Sclose);
        hsize_t d[3] = {0, 0, 0};
        if (H5Sget_simple_extent_ndims(space.id) != 3 || H5Sget_simple_extent_dims(space.id, d, nullptr) != 3)
            return fail(DNSBOX_IO_E_IO, std::string("dataset ") + kNames[f] + " is not three-dimensional");
        if (d[0] != hsize_t(nz) || d[1] != hsize_t(ny) || d[2] != hsize_t(nx))
            return fail(DNSBOX_IO_E_ARG, std::string("dataset ") + kNames[f] + " has Fortran shape (" +
                                             std::to_string(d[2]) + ", " + std::to_string(d[1]) + ", " +
                                             std::to_string(d[0]) + "), expected (" + std::to_string(nx) +
                                             ", " + std::to_string(ny) + ", " + std::to_string(nz) + ")");
        if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, dst[f]) < 0)
            return fail(DNSBOX_IO_E_IO, std::string("cannot read dataset ") + kNames[f]);
    }
    return 0;
}

}  // extern "C"
