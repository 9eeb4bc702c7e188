/* libfkmc_h5.so -- the HDF5 file the reference's program writes (prog/data_save.hxx:32-151,
 * the `output` parameter, default output.h5), over the HDF5 C library.  The driver loads it
 * at run time (dlopen) only when `output` names a .h5 file, so nothing else in the build
 * depends on HDF5.  Plain C, so that the library pulls in no C++ runtime of its own.
 *
 * Every dataset is written by its full path ("mc_data/energies", "stats/cv", ...); missing
 * groups on the path are created.  Doubles are H5T_IEEE_F64LE, integers H5T_STD_I64LE,
 * strings variable-length UTF-8, each in a contiguous layout. */
#include <hdf5.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
    hid_t file, lcpl;
} h5file;

static __thread char g_err[512];

static int fail(const char* what, const char* name) {
    snprintf(g_err, sizeof g_err, "%s%s%s", what, name ? " " : "", name ? name : "");
    return -1;
}

/* dataset `name` of the given types and shape dims[0..rank) (rank 0: scalar) from `data` */
static int write_any(h5file* f, const char* name, hid_t file_type, hid_t mem_type, const void* data, int rank,
                     const uint64_t* dims) {
    hsize_t d[8];
    if (!f || f->file < 0) return fail("archive not open", NULL);
    if (!name || !*name) return fail("empty dataset name", NULL);
    if (rank < 0 || rank > 8) return fail("bad rank for dataset", name);
    for (int i = 0; i < rank; ++i) d[i] = (hsize_t)dims[i];
    const hid_t space = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, d, NULL);
    if (space < 0) return fail("cannot create the dataspace of", name);
    const hid_t ds = H5Dcreate2(f->file, name, file_type, space, f->lcpl, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(space);
        return fail("cannot create dataset", name);
    }
    const herr_t w = H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    H5Sclose(space);
    return w < 0 ? fail("cannot write dataset", name) : 0;
}

/* Creates (truncates) `path`; NULL on failure (fkh5_error() says why). */
void* fkh5_create(const char* path) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL); /* errors come back as return codes */
    const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) {
        fail("cannot create HDF5 file", path);
        return NULL;
    }
    h5file* f = (h5file*)malloc(sizeof(h5file));
    f->file = file;
    f->lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(f->lcpl, 1);
    return f;
}

int fkh5_write_f64(void* h, const char* name, const double* data, int rank, const uint64_t* dims) {
    return write_any((h5file*)h, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, data, rank, dims);
}

int fkh5_write_i64(void* h, const char* name, const int64_t* data, int rank, const uint64_t* dims) {
    return write_any((h5file*)h, name, H5T_STD_I64LE, H5T_NATIVE_INT64, data, rank, dims);
}

int fkh5_write_str(void* h, const char* name, const char* s) {
    const hid_t t = H5Tcopy(H5T_C_S1);
    const char* p[1] = {s};
    H5Tset_size(t, H5T_VARIABLE);
    H5Tset_cset(t, H5T_CSET_UTF8);
    const int r = write_any((h5file*)h, name, t, t, p, 0, NULL);
    H5Tclose(t);
    return r;
}

int fkh5_close(void* h) {
    h5file* f = (h5file*)h;
    int r = 0;
    if (!f) return 0;
    if (f->lcpl >= 0) H5Pclose(f->lcpl);
    if (f->file >= 0 && H5Fclose(f->file) < 0) r = fail("cannot close the HDF5 file", NULL);
    free(f);
    return r;
}

const char* fkh5_error(void) { return g_err; }
