/*
 * h5io.c -- minimal HDF5 access used to write and read the file layout of
 * ClassicalSpinMC.jl's hdf5.jl (the .h5.params metadata file, the per-rank
 * configuration files, IC files).  Plain C over the HDF5 1.10 C library.
 *
 * Arrays are passed in C order with the HDF5 dimensions as they appear on
 * disk (a Julia array's dimensions reversed).  Strings are fixed-length
 * UTF-8, which is what readers of the reference's files (h5py's asstr() and
 * attrs[...].decode()) accept.
 */
#include <hdf5.h>
#include <stdint.h>
#include <string.h>

static int inited = 0;
static void init(void)
{
    if (!inited) {
        H5open();
        H5Eset_auto2(H5E_DEFAULT, NULL, NULL); /* errors are returned, not printed */
        inited = 1;
    }
}

/* mode: 0 read-only, 1 read-write, 2 create (truncate) */
int64_t h5c_open(const char *path, int mode)
{
    init();
    if (mode == 2) return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    return H5Fopen(path, mode == 1 ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
}

int h5c_close(int64_t fid) { return H5Fclose(fid) < 0 ? -1 : 0; }

int h5c_exists(int64_t loc, const char *name)
{
    init();
    return H5Lexists(loc, name, H5P_DEFAULT) > 0;
}

/* open (create = 0) or create (create = 1) a group */
int64_t h5c_group(int64_t loc, const char *name, int create)
{
    init();
    if (create) return H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    return H5Gopen2(loc, name, H5P_DEFAULT);
}

int h5c_gclose(int64_t gid) { return H5Gclose(gid) < 0 ? -1 : 0; }

int h5c_delete(int64_t loc, const char *name)
{
    init();
    return H5Ldelete(loc, name, H5P_DEFAULT) < 0 ? -1 : 0;
}

static hid_t str_type(size_t len)
{
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, len > 0 ? len : 1);
    H5Tset_strpad(t, H5T_STR_NULLPAD);
    H5Tset_cset(t, H5T_CSET_UTF8);
    return t;
}

/* write a dataset (kind 0 float64, 1 int64, 2 uint8); an existing dataset of
 * the same shape is overwritten in place (write_MC_checkpoint), otherwise
 * replaced */
static int write_ds(int64_t loc, const char *name, int kind, int rank, const int64_t *dims,
                    const void *data)
{
    init();
    hsize_t hd[8];
    for (int r = 0; r < rank; ++r) hd[r] = (hsize_t)dims[r];
    hid_t mt = kind == 2 ? H5T_NATIVE_UINT8 : kind ? H5T_NATIVE_INT64 : H5T_NATIVE_DOUBLE;
    hid_t ft = kind == 2 ? H5T_STD_U8LE : kind ? H5T_STD_I64LE : H5T_IEEE_F64LE;
    if (H5Lexists(loc, name, H5P_DEFAULT) > 0) {
        hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
        if (ds >= 0) {
            hid_t sp = H5Dget_space(ds);
            int same = H5Sget_simple_extent_ndims(sp) == rank;
            hsize_t cur[8];
            if (same) {
                H5Sget_simple_extent_dims(sp, cur, NULL);
                for (int r = 0; r < rank; ++r) same &= (cur[r] == hd[r]);
            }
            H5Sclose(sp);
            if (same) {
                herr_t e = H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
                H5Dclose(ds);
                return e < 0 ? -1 : 0;
            }
            H5Dclose(ds);
        }
        H5Ldelete(loc, name, H5P_DEFAULT);
    }
    hid_t sp = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, hd, NULL);
    hid_t ds = H5Dcreate2(loc, name, ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = ds < 0 ? -1 : H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return e < 0 ? -1 : 0;
}

int h5c_write_f64(int64_t loc, const char *name, int rank, const int64_t *dims, const double *d)
{
    return write_ds(loc, name, 0, rank, dims, d);
}

int h5c_write_i64(int64_t loc, const char *name, int rank, const int64_t *dims, const int64_t *d)
{
    return write_ds(loc, name, 1, rank, dims, d);
}

int h5c_write_u8(int64_t loc, const char *name, int rank, const int64_t *dims, const uint8_t *d)
{
    return write_ds(loc, name, 2, rank, dims, d);
}

int h5c_write_str(int64_t loc, const char *name, const char *s)
{
    init();
    if (H5Lexists(loc, name, H5P_DEFAULT) > 0) H5Ldelete(loc, name, H5P_DEFAULT);
    hid_t t = str_type(strlen(s));
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t ds = H5Dcreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = ds < 0 ? -1 : H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, s);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    H5Tclose(t);
    return e < 0 ? -1 : 0;
}

static int write_attr(int64_t loc, const char *name, hid_t ft, hid_t mt, const void *v)
{
    init();
    if (H5Aexists(loc, name) > 0) H5Adelete(loc, name);
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(loc, name, ft, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, mt, v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return e < 0 ? -1 : 0;
}

int h5c_attr_f64(int64_t loc, const char *name, double v)
{
    return write_attr(loc, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &v);
}

int h5c_attr_i64(int64_t loc, const char *name, int64_t v)
{
    return write_attr(loc, name, H5T_STD_I64LE, H5T_NATIVE_INT64, &v);
}

int h5c_attr_str(int64_t loc, const char *name, const char *s)
{
    init();
    hid_t t = str_type(strlen(s));
    int r = write_attr(loc, name, t, t, s);
    H5Tclose(t);
    return r;
}

/* class of a dataset: 0 float, 1 integer, 2 string, 3 unsigned 8-bit
 * integer, -1 missing; rank and dims (up to 8) on disk; for strings dims[0] =
 * the string size */
int h5c_info(int64_t loc, const char *name, int *kind, int *rank, int64_t *dims)
{
    init();
    if (H5Lexists(loc, name, H5P_DEFAULT) <= 0) return -1;
    hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    hid_t t = H5Dget_type(ds);
    H5T_class_t c = H5Tget_class(t);
    hid_t sp = H5Dget_space(ds);
    int n = H5Sget_simple_extent_ndims(sp);
    hsize_t hd[8];
    H5Sget_simple_extent_dims(sp, hd, NULL);
    *rank = n;
    for (int r = 0; r < n && r < 8; ++r) dims[r] = (int64_t)hd[r];
    if (c == H5T_STRING) {
        *kind = 2;
        dims[0] = H5Tis_variable_str(t) ? -1 : (int64_t)H5Tget_size(t);
    } else if (c == H5T_INTEGER && H5Tget_size(t) == 1 && H5Tget_sign(t) == H5T_SGN_NONE) {
        *kind = 3;
    } else {
        *kind = c == H5T_INTEGER ? 1 : 0;
    }
    H5Sclose(sp);
    H5Tclose(t);
    H5Dclose(ds);
    return 0;
}

static int read_ds(int64_t loc, const char *name, hid_t mt, void *out)
{
    init();
    hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    herr_t e = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(ds);
    return e < 0 ? -1 : 0;
}

int h5c_read_f64(int64_t loc, const char *name, double *out)
{
    return read_ds(loc, name, H5T_NATIVE_DOUBLE, out);
}

int h5c_read_i64(int64_t loc, const char *name, int64_t *out)
{
    return read_ds(loc, name, H5T_NATIVE_INT64, out);
}

int h5c_read_u8(int64_t loc, const char *name, uint8_t *out)
{
    return read_ds(loc, name, H5T_NATIVE_UINT8, out);
}

/* string dataset (fixed or variable length) into buf (NUL-terminated) */
int h5c_read_str(int64_t loc, const char *name, char *buf, int buflen)
{
    init();
    hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    hid_t t = H5Dget_type(ds);
    herr_t e;
    if (H5Tis_variable_str(t)) {
        char *p = NULL;
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        e = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p);
        if (e >= 0 && p) {
            strncpy(buf, p, (size_t)buflen - 1);
            buf[buflen - 1] = 0;
            H5free_memory(p);
        }
        H5Tclose(mt);
    } else {
        size_t n = H5Tget_size(t);
        if ((int)n >= buflen) n = (size_t)buflen - 1;
        memset(buf, 0, (size_t)buflen);
        hid_t mt = str_type(n);
        e = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        H5Tclose(mt);
    }
    H5Tclose(t);
    H5Dclose(ds);
    return e < 0 ? -1 : 0;
}

/* attribute: kind 0 float, 1 int, 2 string; value in *f / *i / buf */
int h5c_attr_read(int64_t loc, const char *name, int *kind, double *f, int64_t *i, char *buf,
                  int buflen)
{
    init();
    if (H5Aexists(loc, name) <= 0) return -1;
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    hid_t t = H5Aget_type(a);
    H5T_class_t c = H5Tget_class(t);
    herr_t e = -1;
    if (c == H5T_STRING) {
        *kind = 2;
        memset(buf, 0, (size_t)buflen);
        if (H5Tis_variable_str(t)) {
            char *p = NULL;
            hid_t mt = H5Tcopy(H5T_C_S1);
            H5Tset_size(mt, H5T_VARIABLE);
            e = H5Aread(a, mt, &p);
            if (e >= 0 && p) {
                strncpy(buf, p, (size_t)buflen - 1);
                H5free_memory(p);
            }
            H5Tclose(mt);
        } else {
            size_t n = H5Tget_size(t);
            if ((int)n >= buflen) n = (size_t)buflen - 1;
            hid_t mt = str_type(n);
            e = H5Aread(a, mt, buf);
            H5Tclose(mt);
        }
    } else if (c == H5T_INTEGER) {
        *kind = 1;
        e = H5Aread(a, H5T_NATIVE_INT64, i);
    } else {
        *kind = 0;
        e = H5Aread(a, H5T_NATIVE_DOUBLE, f);
    }
    H5Tclose(t);
    H5Aclose(a);
    return e < 0 ? -1 : 0;
}

/* newline-separated names of the links (idx_type 0) or attributes
 * (idx_type 1) of loc, in creation-independent name order */
struct lst {
    char *buf;
    int len, pos;
};

static herr_t add_name(const char *name, struct lst *l)
{
    int n = (int)strlen(name);
    if (l->pos + n + 1 < l->len) {
        memcpy(l->buf + l->pos, name, (size_t)n);
        l->buf[l->pos + n] = '\n';
        l->pos += n + 1;
        l->buf[l->pos] = 0;
    }
    return 0;
}

static herr_t link_cb(hid_t g, const char *name, const H5L_info_t *info, void *op)
{
    (void)g;
    (void)info;
    return add_name(name, (struct lst *)op);
}

static herr_t attr_cb(hid_t g, const char *name, const H5A_info_t *info, void *op)
{
    (void)g;
    (void)info;
    return add_name(name, (struct lst *)op);
}

int h5c_list(int64_t loc, int what, char *buf, int buflen)
{
    init();
    struct lst l = {buf, buflen, 0};
    buf[0] = 0;
    herr_t e;
    if (what == 0) e = H5Literate(loc, H5_INDEX_NAME, H5_ITER_INC, NULL, link_cb, &l);
    else e = H5Aiterate2(loc, H5_INDEX_NAME, H5_ITER_INC, NULL, attr_cb, &l);
    return e < 0 ? -1 : l.pos;
}
