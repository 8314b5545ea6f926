// HDF5 file access for the particle dump (output:dump_all_particles: hdf5) and
// the beam import (beam:from_hdf5), over libhdf5's C API. Host code only: the
// particles are fetched from the device first.
//
// The file layout follows the reference's writer (hdf5-writer/src/data.rs,
// dataset.rs, datatype.rs): scalars in a scalar dataspace; 1-D datasets of
// scalars; four-vectors and Stokes vectors as 1-D datasets of a 4-element
// array type of IEEE f64; strings as fixed-length, NUL-terminated UTF-8 C
// strings of their byte length; booleans and enumerations as enumerations over
// u8 ("false"/"true", "linear"/"circular", "cos^2"/"flattop"/"gaussian"); the
// `unit` and `desc` of a dataset as fixed-length string attributes; an alias as
// a soft link beside the dataset.
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

struct Quiet {  // libhdf5 prints its error stack by default; errors are returned instead
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};
const Quiet quiet;

// Splits "a/b/c" into the parent group path "a/b" and the leaf "c".
void split(const char* path, std::string& parent, std::string& leaf) {
    std::string p(path);
    const size_t k = p.rfind('/');
    if (k == std::string::npos) {
        parent = ".";
        leaf = p;
    } else {
        parent = k == 0 ? "/" : p.substr(0, k);
        leaf = p.substr(k + 1);
    }
}

// Opens (creating as needed) every group along `path`.
hid_t open_groups(hid_t f, const std::string& path) {
    if (path == "." || path == "/") return H5Gopen2(f, "/", H5P_DEFAULT);
    hid_t g = H5Gopen2(f, "/", H5P_DEFAULT);
    size_t start = path[0] == '/' ? 1 : 0;
    while (g >= 0 && start < path.size()) {
        size_t end = path.find('/', start);
        if (end == std::string::npos) end = path.size();
        const std::string name = path.substr(start, end - start);
        hid_t next = H5Lexists(g, name.c_str(), H5P_DEFAULT) > 0 ? H5Gopen2(g, name.c_str(), H5P_DEFAULT)
                                                                   : H5Gcreate2(g, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5Gclose(g);
        g = next;
        start = end + 1;
    }
    return g;
}

hid_t string_type(size_t len) {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, len > 0 ? len : 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    H5Tset_cset(t, H5T_CSET_UTF8);
    return t;
}

int attach(hid_t d, const char* name, const char* value) {
    if (!value) return 0;
    const size_t len = std::strlen(value);
    hid_t t = string_type(len);
    hid_t s = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(d, name, t, s, H5P_DEFAULT, H5P_DEFAULT);
    herr_t rc = a >= 0 ? H5Awrite(a, t, value) : -1;
    if (a >= 0) H5Aclose(a);
    H5Sclose(s);
    H5Tclose(t);
    return rc < 0 ? fail(std::string("cannot write attribute ") + name) : 0;
}

// Element types of the writer/reader calls.
enum Kind : int { K_F64 = 0, K_U64 = 1, K_I64 = 2, K_BOOL = 3, K_U8ENUM = 4 };

hid_t element_type(int kind, int width, const char* const* names, const uint8_t* values, int n_names) {
    hid_t base;
    switch (kind) {
    case K_F64: base = H5Tcopy(H5T_IEEE_F64LE); break;
    case K_U64: base = H5Tcopy(H5T_STD_U64LE); break;
    case K_I64: base = H5Tcopy(H5T_STD_I64LE); break;
    case K_BOOL: {
        base = H5Tenum_create(H5T_STD_U8LE);
        const uint8_t f = 0, t = 1;
        H5Tenum_insert(base, "false", &f);
        H5Tenum_insert(base, "true", &t);
        break;
    }
    case K_U8ENUM: {
        base = H5Tenum_create(H5T_STD_U8LE);
        for (int k = 0; k < n_names; ++k) H5Tenum_insert(base, names[k], &values[k]);
        break;
    }
    default: return -1;
    }
    if (width <= 1) return base;
    const hsize_t dims[1] = {hsize_t(width)};
    hid_t arr = H5Tarray_create2(base, 1, dims);
    H5Tclose(base);
    return arr;
}

int write_dataset(hid_t f, const char* path, hid_t type, hid_t mem_type, hid_t space, const void* data, const char* unit,
                  const char* desc) {
    std::string parent, leaf;
    split(path, parent, leaf);
    hid_t g = open_groups(f, parent);
    if (g < 0) return fail(std::string("cannot open group ") + parent);
    hid_t d = H5Dcreate2(g, leaf.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Gclose(g);
    if (d < 0) return fail(std::string("cannot create dataset ") + path);
    herr_t rc = H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    int out = rc < 0 ? fail(std::string("cannot write dataset ") + path) : 0;
    if (!out) out = attach(d, "unit", unit);
    if (!out) out = attach(d, "desc", desc);
    H5Dclose(d);
    return out;
}

}  // namespace

extern "C" {

const char* ptg_h5_last_error() { return g_err.c_str(); }

// A new file (truncating an existing one); returns its handle, or < 0.
int64_t ptg_h5_create(const char* path) {
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot create HDF5 file ") + path);
    return f;
}

// An existing file, read-only; returns its handle, or < 0.
int64_t ptg_h5_open(const char* path) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot open HDF5 file ") + path);
    return f;
}

int ptg_h5_close(int64_t f) { return H5Fclose(hid_t(f)) < 0 ? fail("cannot close HDF5 file") : 0; }

// An empty group (and its parents).
int ptg_h5_group(int64_t f, const char* path) {
    hid_t g = open_groups(hid_t(f), path);
    if (g < 0) return fail(std::string("cannot create group ") + path);
    H5Gclose(g);
    return 0;
}

// A scalar dataset of `kind` (K_U8ENUM takes the enumeration's n_names names and values).
int ptg_h5_write_scalar(int64_t f, const char* path, int kind, const void* value, const char* const* names,
                        const uint8_t* values, int n_names, const char* unit, const char* desc) {
    hid_t t = element_type(kind, 1, names, values, n_names);
    if (t < 0) return fail("unknown element kind");
    hid_t s = H5Screate(H5S_SCALAR);
    int rc = write_dataset(hid_t(f), path, t, t, s, value, unit, desc);
    H5Sclose(s);
    H5Tclose(t);
    return rc;
}

// A scalar string dataset.
int ptg_h5_write_string(int64_t f, const char* path, const char* value, const char* unit, const char* desc) {
    hid_t t = string_type(std::strlen(value));
    hid_t s = H5Screate(H5S_SCALAR);
    int rc = write_dataset(hid_t(f), path, t, t, s, value, unit, desc);
    H5Sclose(s);
    H5Tclose(t);
    return rc;
}

// A 1-D dataset of n elements; width 4 makes each element a 4-element array
// (four-vectors, Stokes vectors), read from n x 4 row-major data.
int ptg_h5_write_array(int64_t f, const char* path, int kind, const void* data, uint64_t n, int width, const char* unit,
                       const char* desc) {
    hid_t t = element_type(kind, width, nullptr, nullptr, 0);
    if (t < 0) return fail("unknown element kind");
    const hsize_t dims[1] = {hsize_t(n)};
    hid_t s = H5Screate_simple(1, dims, nullptr);
    static const double dummy = 0.0;
    int rc = write_dataset(hid_t(f), path, t, t, s, n ? data : &dummy, unit, desc);
    H5Sclose(s);
    H5Tclose(t);
    return rc;
}

// A soft link `alias` to `target`, both names inside `group`.
int ptg_h5_alias(int64_t f, const char* group, const char* target, const char* alias) {
    hid_t g = open_groups(hid_t(f), group);
    if (g < 0) return fail(std::string("cannot open group ") + group);
    herr_t rc = H5Lcreate_soft(target, g, alias, H5P_DEFAULT, H5P_DEFAULT);
    H5Gclose(g);
    return rc < 0 ? fail(std::string("cannot create alias ") + alias) : 0;
}

// Whether `path` names an object (1), not (0), or an error (< 0).
int ptg_h5_exists(int64_t f, const char* path) {
    std::string p(path);
    size_t start = p[0] == '/' ? 1 : 0;
    while (start < p.size()) {  // H5Lexists needs every parent to exist
        size_t end = p.find('/', start);
        if (end == std::string::npos) end = p.size();
        if (H5Lexists(hid_t(f), p.substr(0, end).c_str(), H5P_DEFAULT) <= 0) return 0;
        start = end + 1;
    }
    return H5Oexists_by_name(hid_t(f), path, H5P_DEFAULT) > 0 ? 1 : 0;
}

// Shape of a dataset: element count (1 for a scalar), array width (1 when the
// element is not an array), and the element's type class (H5T_class_t of the
// array's base type).
int ptg_h5_info(int64_t f, const char* path, uint64_t* n, int* width, int* cls) {
    hid_t d = H5Dopen2(hid_t(f), path, H5P_DEFAULT);
    if (d < 0) return fail(std::string("no dataset ") + path);
    hid_t s = H5Dget_space(d);
    hid_t t = H5Dget_type(d);
    *n = uint64_t(H5Sget_simple_extent_npoints(s));
    *width = 1;
    H5T_class_t c = H5Tget_class(t);
    if (c == H5T_ARRAY) {
        hsize_t dims[4] = {0, 0, 0, 0};
        const int nd = H5Tget_array_ndims(t);
        H5Tget_array_dims2(t, dims);
        *width = nd == 1 ? int(dims[0]) : -1;
        hid_t b = H5Tget_super(t);
        c = H5Tget_class(b);
        H5Tclose(b);
    }
    *cls = int(c);
    H5Tclose(t);
    H5Sclose(s);
    H5Dclose(d);
    return 0;
}

// Reads a dataset of n elements of the given width as f64 (kind 0) or u64
// (kind 1), converting from the stored numeric type; booleans and other u8
// enumerations read as their integer values (kind 1).
int ptg_h5_read(int64_t f, const char* path, int kind, void* out, uint64_t n, int width) {
    uint64_t n0;
    int w0, cls;
    if (ptg_h5_info(f, path, &n0, &w0, &cls)) return -1;
    if (n0 != n || w0 != width) return fail(std::string("unexpected shape of ") + path);
    hid_t d = H5Dopen2(hid_t(f), path, H5P_DEFAULT);
    hid_t mt;
    if (cls == H5T_ENUM) {  // read the enumeration through its own type, then widen
        hid_t ft = H5Dget_type(d);
        std::vector<uint8_t> raw(n * size_t(width));
        herr_t rc = H5Dread(d, ft, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw.data());
        H5Tclose(ft);
        H5Dclose(d);
        if (rc < 0) return fail(std::string("cannot read ") + path);
        for (size_t k = 0; k < raw.size(); ++k) {
            if (kind == K_F64) static_cast<double*>(out)[k] = raw[k];
            else static_cast<uint64_t*>(out)[k] = raw[k];
        }
        return 0;
    }
    hid_t base = kind == K_F64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_UINT64;
    if (width > 1) {
        const hsize_t dims[1] = {hsize_t(width)};
        mt = H5Tarray_create2(base, 1, dims);
    } else {
        mt = H5Tcopy(base);
    }
    herr_t rc = H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Tclose(mt);
    H5Dclose(d);
    return rc < 0 ? fail(std::string("cannot read ") + path) : 0;
}

// Reads a scalar string dataset (fixed or variable length) into out[cap].
int ptg_h5_read_string(int64_t f, const char* path, char* out, int cap) {
    hid_t d = H5Dopen2(hid_t(f), path, H5P_DEFAULT);
    if (d < 0) return fail(std::string("no dataset ") + path);
    hid_t t = H5Dget_type(d);
    int rc = 0;
    if (H5Tget_class(t) != H5T_STRING) {
        rc = fail(std::string("not a string: ") + path);
    } else if (H5Tis_variable_str(t) > 0) {
        char* s = nullptr;
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s) < 0 || !s) {
            rc = fail(std::string("cannot read ") + path);
        } else {
            std::strncpy(out, s, size_t(cap) - 1);
            out[cap - 1] = 0;
            H5free_memory(s);
        }
        H5Tclose(mt);
    } else {
        const size_t len = H5Tget_size(t);
        std::vector<char> buf(len + 1, 0);
        if (H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) {
            rc = fail(std::string("cannot read ") + path);
        } else {
            std::strncpy(out, buf.data(), size_t(cap) - 1);
            out[cap - 1] = 0;
        }
    }
    H5Tclose(t);
    H5Dclose(d);
    return rc;
}

// Reads a string attribute of a dataset into out[cap] ("" if absent).
int ptg_h5_read_attribute(int64_t f, const char* path, const char* name, char* out, int cap) {
    out[0] = 0;
    hid_t d = H5Oopen(hid_t(f), path, H5P_DEFAULT);
    if (d < 0) return fail(std::string("no object ") + path);
    int rc = 0;
    if (H5Aexists(d, name) > 0) {
        hid_t a = H5Aopen(d, name, H5P_DEFAULT);
        hid_t t = H5Aget_type(a);
        const size_t len = H5Tget_size(t);
        std::vector<char> buf(len + 1, 0);
        if (H5Tis_variable_str(t) > 0 || H5Aread(a, t, buf.data()) < 0) rc = fail(std::string("cannot read attribute ") + name);
        else {
            std::strncpy(out, buf.data(), size_t(cap) - 1);
            out[cap - 1] = 0;
        }
        H5Tclose(t);
        H5Aclose(a);
    }
    H5Oclose(d);
    return rc;
}

}  // extern "C"
