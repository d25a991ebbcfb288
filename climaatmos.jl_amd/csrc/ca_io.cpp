// ClimaCore HDF5 checkpoint I/O (host side): the file format of ClimaCore
// 0.15.1's InputOutput.HDF5Writer / HDF5Reader as ClimaAtmos uses it in
// save_state_to_disk_func (src/callbacks/callbacks.jl:297-320) and
// get_state_restart (src/simulation/restart.jl:22-43).  Built with g++
// against the image's libhdf5 (1.10); files use the oldest on-disk format
// versions, which every libhdf5 from 1.8 on reads.
#include <hdf5.h>

#include <cstring>
#include <string>
#include <vector>

#include "caio.h"

namespace {

thread_local std::string g_err;

ca_status fail(ca_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

// Silences libhdf5's error printing for the duration of one entry point.
struct Quiet {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// Closes an HDF5 identifier when it goes out of scope.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Id(const Id &) = delete;
    Id &operator=(const Id &) = delete;
    Id &operator=(Id &&o) noexcept {
        reset();
        id = o.id;
        close = o.close;
        o.id = -1;
        return *this;
    }
    ~Id() { reset(); }
    void reset() {
        if (id >= 0 && close) close(id);
        id = -1;
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

struct H5Error {
    std::string what;
};

hid_t must(hid_t id, const std::string &what) {
    if (id < 0) throw H5Error{what};
    return id;
}
void must(herr_t st, const char *what, const std::string &name) {
    if (st < 0) throw H5Error{std::string(what) + " " + name};
}

// ---------------------------------------------------------------- writing
// Fixed-length UTF-8 string, null-terminated, as HDF5.jl stores a String.
void attr_str(hid_t obj, const char *name, const std::string &v) {
    Id t(must(H5Tcopy(H5T_C_S1), "H5Tcopy"), H5Tclose);
    must(H5Tset_size(t, v.size() + 1), "H5Tset_size", name);
    must(H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset", name);
    must(H5Tset_strpad(t, H5T_STR_NULLTERM), "H5Tset_strpad", name);
    Id sp(must(H5Screate(H5S_SCALAR), "H5Screate"), H5Sclose);
    Id a(must(H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), std::string("H5Acreate2 ") + name),
         H5Aclose);
    must(H5Awrite(a, t, v.c_str()), "H5Awrite", name);
}

// Vector{String}: a 1-d attribute of variable-length UTF-8 strings.
void attr_strs(hid_t obj, const char *name, const std::vector<std::string> &v) {
    Id t(must(H5Tcopy(H5T_C_S1), "H5Tcopy"), H5Tclose);
    must(H5Tset_size(t, H5T_VARIABLE), "H5Tset_size", name);
    must(H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset", name);
    hsize_t n = v.size();
    Id sp(must(H5Screate_simple(1, &n, nullptr), "H5Screate_simple"), H5Sclose);
    Id a(must(H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), std::string("H5Acreate2 ") + name),
         H5Aclose);
    std::vector<const char *> p;
    for (auto &s : v) p.push_back(s.c_str());
    must(H5Awrite(a, t, p.data()), "H5Awrite", name);
}

template <class T>
void attr_num(hid_t obj, const char *name, hid_t file_type, hid_t mem_type, const T *v, hsize_t n,
              bool scalar) {
    Id sp(must(scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &n, nullptr), "H5Screate"), H5Sclose);
    Id a(must(H5Acreate2(obj, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT), std::string("H5Acreate2 ") + name),
         H5Aclose);
    must(H5Awrite(a, mem_type, v), "H5Awrite", name);
}
void attr_i64(hid_t obj, const char *name, int64_t v) {
    attr_num(obj, name, H5T_STD_I64LE, H5T_NATIVE_INT64, &v, 1, true);
}
void attr_f32(hid_t obj, const char *name, float v) {
    attr_num(obj, name, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &v, 1, true);
}

Id group(hid_t file, const std::string &path, hid_t lcpl) {
    return Id(must(H5Gcreate2(file, path.c_str(), lcpl, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2 " + path),
              H5Gclose);
}

void dataset_f32(hid_t file, const std::string &path, hid_t lcpl, const std::vector<hsize_t> &dims,
                 const float *data, const char *layout, const char *value_type, const char *grid,
                 const char *staggering) {
    Id sp(must(H5Screate_simple((int)dims.size(), dims.data(), nullptr), "H5Screate_simple"), H5Sclose);
    Id d(must(H5Dcreate2(file, path.c_str(), H5T_IEEE_F32LE, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT),
              "H5Dcreate2 " + path),
         H5Dclose);
    must(H5Dwrite(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "H5Dwrite", path);
    attr_str(d, "type", "Field");
    attr_str(d, "data_layout", layout);
    attr_str(d, "value_type", value_type);
    attr_str(d, "grid", grid);
    if (staggering) attr_str(d, "staggering", staggering);
}

// value_type as Julia 1.10 and later print the element type of Y.c and Y.f
// (ClimaCore's reader evaluates the string, so either NamedTuple spelling works)
const char *kCentreType =
    "@NamedTuple{\xCF\x81::Float32, u\xE2\x82\x95::ClimaCore.Geometry.Covariant12Vector{Float32}, "
    "\xCF\x81" "e_tot::Float32}";
const char *kFaceType = "@NamedTuple{u\xE2\x82\x83::ClimaCore.Geometry.Covariant3Vector{Float32}}";

void write_file(const char *path, const ca_io_grid *g, const float *c, const float *f, double time,
                uint64_t hash) {
    Id fcpl(must(H5Pcreate(H5P_FILE_CREATE), "H5Pcreate"), H5Pclose);
    Id file(must(H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), std::string("H5Fcreate ") + path), H5Fclose);
    Id lcpl(must(H5Pcreate(H5P_LINK_CREATE), "H5Pcreate"), H5Pclose);
    must(H5Pset_create_intermediate_group(lcpl, 1), "H5Pset_create_intermediate_group", "");
    must(H5Pset_char_encoding(lcpl, H5T_CSET_UTF8), "H5Pset_char_encoding", "");

    // root attributes: HDF5Writer's version stamp, then the two ClimaAtmos adds
    attr_str(file, "ClimaCore version", "0.15.1");
    attr_f32(file, "time", (float)time);
    attr_num(file, "atmos_model_hash", H5T_STD_U64LE, H5T_NATIVE_UINT64, &hash, 1, true);

    const int nl = g->nlev;
    std::vector<float> faces(nl + 1);
    for (int k = 0; k <= nl; ++k) faces[k] = (float)g->z_face[k];
    {
        Id d = group(file, "domains/z", lcpl);
        attr_str(d, "type", "IntervalDomain");
        attr_str(d, "coord_type", "ZPoint");
        attr_f32(d, "coord_min", faces[0]);
        attr_f32(d, "coord_max", faces[nl]);
        attr_strs(d, "boundary_names", {"bottom", "top"});
        Id m = group(file, "meshes/z", lcpl);
        attr_str(m, "type", "IntervalMesh");
        attr_str(m, "domain", "z");
        attr_i64(m, "nelements", nl);
        attr_str(m, "faces_type", "Array");
        attr_num(m.id, "faces", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, faces.data(), (hsize_t)(nl + 1), false);
        Id t = group(file, "topologies/z", lcpl);
        attr_str(t, "type", "IntervalTopology");
        attr_str(t, "mesh", "z");
        Id v = group(file, "grids/vertical_grid", lcpl);
        attr_str(v, "type", "FiniteDifferenceGrid");
        attr_str(v, "topology", "z");
    }
    const char *grid_name = "vertical_grid";
    if (g->kind == CA_IO_SPHERE) {
        const int64_t nh = 6ll * g->h_elem * g->h_elem;
        Id d = group(file, "domains/sphere", lcpl);
        attr_str(d, "type", "SphereDomain");
        attr_f32(d, "radius", (float)g->radius);
        Id m = group(file, "meshes/cubedsphere", lcpl);
        attr_str(m, "type", "EquiangularCubedSphere");
        attr_i64(m, "ne", g->h_elem);
        attr_str(m, "localelementmap", "NormalizedBilinearMap");
        attr_str(m, "domain", "sphere");
        Id t = group(file, "topologies/2d", lcpl);
        attr_str(t, "type", "Topology2D");
        attr_str(t, "mesh", "cubedsphere");
        if (g->elemorder) {
            // Julia's vec of the 3 x Nh matrix of CartesianIndex components
            std::vector<int64_t> eo(3 * nh);
            for (int64_t i = 0; i < 3 * nh; ++i) eo[i] = g->elemorder[i];
            hsize_t dims[2] = {(hsize_t)nh, 3};
            Id sp(must(H5Screate_simple(2, dims, nullptr), "H5Screate_simple"), H5Sclose);
            Id a(must(H5Acreate2(t, "elemorder", H5T_STD_I64LE, sp, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2 elemorder"),
                 H5Aclose);
            must(H5Awrite(a, H5T_NATIVE_INT64, eo.data()), "H5Awrite", "elemorder");
        }
        Id hg = group(file, "grids/horizontal_grid", lcpl);
        attr_str(hg, "type", "SpectralElementGrid2D");
        attr_str(hg, "quadrature_type", "GLL");
        attr_i64(hg, "quadrature_num_points", g->nq);
        attr_str(hg, "topology", "2d");
        Id eg = group(file, "grids/extruded_finite_difference_grid", lcpl);
        attr_str(eg, "type", "ExtrudedFiniteDifferenceGrid");
        attr_str(eg, "horizontal_grid", "horizontal_grid");
        attr_str(eg, "vertical_grid", "vertical_grid");
        attr_str(eg, "hypsography_type", "Flat");
        attr_str(eg, "global_geometry", g->deep ? "DeepSphericalGlobalGeometry" : "ShallowSphericalGlobalGeometry");
        grid_name = "extruded_finite_difference_grid";
    }

    Id y = group(file, "fields/Y", lcpl);
    attr_str(y, "type", "FieldVector");
    attr_strs(y, "fields", {"c", "f"});

    if (g->kind == CA_IO_COLUMN) {
        dataset_f32(file, "fields/Y.c", lcpl, {4, (hsize_t)nl}, c, "VF", kCentreType, grid_name, "CellCenter");
        dataset_f32(file, "fields/Y.f", lcpl, {1, (hsize_t)(nl + 1)}, f, "VF", kFaceType, grid_name, "CellFace");
        return;
    }
    const hsize_t nq = g->nq, nh = 6ull * g->h_elem * g->h_elem;
    if (g->layout == CA_IO_VIJFH) {
        dataset_f32(file, "fields/Y.c", lcpl, {nh, 4, nq, nq, (hsize_t)nl}, c, "VIJFH", kCentreType, grid_name,
                    "CellCenter");
        dataset_f32(file, "fields/Y.f", lcpl, {nh, 1, nq, nq, (hsize_t)(nl + 1)}, f, "VIJFH", kFaceType, grid_name,
                    "CellFace");
        return;
    }
    // VIJHF: Julia (Nv, Ni, Nj, Nh, F) from the VIJFH input (Nv, Ni, Nj, F, Nh)
    auto to_hf = [&](const float *src, int F, int nv) {
        const size_t blk = (size_t)nv * nq * nq;
        std::vector<float> out(blk * F * nh);
        for (hsize_t h = 0; h < nh; ++h)
            for (int q = 0; q < F; ++q)
                std::memcpy(&out[(q * nh + h) * blk], &src[(h * F + q) * blk], blk * sizeof(float));
        return out;
    };
    auto cc = to_hf(c, 4, nl), ff = to_hf(f, 1, nl + 1);
    dataset_f32(file, "fields/Y.c", lcpl, {4, nh, nq, nq, (hsize_t)nl}, cc.data(), "VIJHF", kCentreType, grid_name,
                "CellCenter");
    dataset_f32(file, "fields/Y.f", lcpl, {1, nh, nq, nq, (hsize_t)(nl + 1)}, ff.data(), "VIJHF", kFaceType,
                grid_name, "CellFace");
}

// ---------------------------------------------------------------- reading
std::string read_str_attr(hid_t obj, const char *name) {
    Id a(must(H5Aopen(obj, name, H5P_DEFAULT), std::string("missing attribute ") + name), H5Aclose);
    Id t(must(H5Aget_type(a), "H5Aget_type"), H5Tclose);
    if (H5Tget_class(t) != H5T_STRING) throw H5Error{std::string("attribute ") + name + " is not a string"};
    Id sp(must(H5Aget_space(a), "H5Aget_space"), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp) != 1) throw H5Error{std::string("attribute ") + name + " is not a scalar"};
    if (H5Tis_variable_str(t) > 0) {
        Id mt(must(H5Tcopy(H5T_C_S1), "H5Tcopy"), H5Tclose);
        must(H5Tset_size(mt, H5T_VARIABLE), "H5Tset_size", name);
        must(H5Tset_cset(mt, H5Tget_cset(t)), "H5Tset_cset", name);  // libhdf5 does not convert charsets
        char *p = nullptr;
        must(H5Aread(a, mt, &p), "H5Aread", name);
        std::string s = p ? p : "";
        H5free_memory(p);
        return s;
    }
    size_t n = H5Tget_size(t);
    std::vector<char> buf(n + 1, 0);
    must(H5Aread(a, t, buf.data()), "H5Aread", name);
    return std::string(buf.data(), strnlen(buf.data(), n));
}

std::vector<std::string> read_strs_attr(hid_t obj, const char *name) {
    Id a(must(H5Aopen(obj, name, H5P_DEFAULT), std::string("missing attribute ") + name), H5Aclose);
    Id t(must(H5Aget_type(a), "H5Aget_type"), H5Tclose);
    if (H5Tget_class(t) != H5T_STRING) throw H5Error{std::string("attribute ") + name + " is not a string"};
    Id sp(must(H5Aget_space(a), "H5Aget_space"), H5Sclose);
    hssize_t n = H5Sget_simple_extent_npoints(sp);
    std::vector<std::string> out;
    if (H5Tis_variable_str(t) > 0) {
        Id mt(must(H5Tcopy(H5T_C_S1), "H5Tcopy"), H5Tclose);
        must(H5Tset_size(mt, H5T_VARIABLE), "H5Tset_size", name);
        must(H5Tset_cset(mt, H5Tget_cset(t)), "H5Tset_cset", name);  // libhdf5 does not convert charsets
        std::vector<char *> p(n, nullptr);
        must(H5Aread(a, mt, p.data()), "H5Aread", name);
        for (auto *s : p) out.emplace_back(s ? s : "");
        H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, p.data());
    } else {
        size_t len = H5Tget_size(t);
        std::vector<char> buf(len * n + 1, 0);
        must(H5Aread(a, t, buf.data()), "H5Aread", name);
        for (hssize_t i = 0; i < n; ++i) out.emplace_back(&buf[i * len], strnlen(&buf[i * len], len));
    }
    return out;
}

double read_num_attr(hid_t obj, const char *name) {
    Id a(must(H5Aopen(obj, name, H5P_DEFAULT), std::string("missing attribute ") + name), H5Aclose);
    double v = 0;
    must(H5Aread(a, H5T_NATIVE_DOUBLE, &v), "H5Aread", name);
    return v;
}

struct FileInfo {
    ca_io_info info{};
    std::vector<hsize_t> cdims, fdims;
};

// Field names of a NamedTuple type string, in either spelling:
// "@NamedTuple{a::T, b::U}" or "NamedTuple{(:a, :b), Tuple{T, U}}".
std::vector<std::string> namedtuple_fields(const std::string &vt) {
    std::vector<std::string> out;
    auto trim = [](std::string x) {
        size_t a = x.find_first_not_of(' '), b = x.find_last_not_of(' ');
        return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
    };
    if (vt.rfind("@NamedTuple{", 0) == 0) {
        int depth = 0;
        std::string cur;
        for (size_t i = 12; i < vt.size(); ++i) {
            char ch = vt[i];
            if ((ch == ',' || ch == '}') && depth == 0) {
                size_t k = cur.find("::");
                out.push_back(trim(k == std::string::npos ? cur : cur.substr(0, k)));
                cur.clear();
                if (ch == '}') break;
                continue;
            }
            if (ch == '{') ++depth;
            if (ch == '}') --depth;
            cur += ch;
        }
    } else if (vt.rfind("NamedTuple{(", 0) == 0) {
        size_t end = vt.find(')');
        if (end == std::string::npos) return out;
        std::string body = vt.substr(12, end - 12), cur;
        for (size_t i = 0; i <= body.size(); ++i) {
            if (i == body.size() || body[i] == ',') {
                std::string x = trim(cur);
                if (!x.empty() && x[0] == ':') out.push_back(x.substr(1));
                cur.clear();
            } else {
                cur += body[i];
            }
        }
    }
    return out;
}

bool names_dry_state(const std::string &vt, bool face) {
    auto names = namedtuple_fields(vt);
    if (face) return names == std::vector<std::string>{"u\xE2\x82\x83"} && vt.find("Covariant3Vector") != std::string::npos;
    return names == std::vector<std::string>{"\xCF\x81", "u\xE2\x82\x95", "\xCF\x81" "e_tot"} &&
           vt.find("Covariant12Vector") != std::string::npos;
}

void inspect(hid_t file, FileInfo &fi) {
    ca_io_info &in = fi.info;
    in.time = read_num_attr(file, "time");
    in.has_hash = H5Aexists(file, "atmos_model_hash") > 0;
    if (in.has_hash) {
        Id a(must(H5Aopen(file, "atmos_model_hash", H5P_DEFAULT), "H5Aopen"), H5Aclose);
        must(H5Aread(a, H5T_NATIVE_UINT64, &in.atmos_model_hash), "H5Aread", "atmos_model_hash");
    }
    if (H5Lexists(file, "fields", H5P_DEFAULT) <= 0 || H5Lexists(file, "fields/Y", H5P_DEFAULT) <= 0)
        throw H5Error{"no fields/Y in the file"};
    {
        Id y(must(H5Oopen(file, "fields/Y", H5P_DEFAULT), "H5Oopen fields/Y"), H5Oclose);
        if (read_str_attr(y, "type") != "FieldVector") throw H5Error{"fields/Y is not a FieldVector"};
        auto names = read_strs_attr(y, "fields");
        if (names.size() != 2 || names[0] != "c" || names[1] != "f")
            throw H5Error{"fields/Y holds other components than (c, f)"};
    }
    std::string lay[2], grid;
    for (int w = 0; w < 2; ++w) {
        const char *nm = w ? "fields/Y.f" : "fields/Y.c";
        if (H5Lexists(file, nm, H5P_DEFAULT) <= 0) throw H5Error{std::string("missing dataset ") + nm};
        Id d(must(H5Dopen2(file, nm, H5P_DEFAULT), std::string("H5Dopen2 ") + nm), H5Dclose);
        if (read_str_attr(d, "type") != "Field") throw H5Error{std::string(nm) + " is not a Field"};
        lay[w] = read_str_attr(d, "data_layout");
        std::string vt = read_str_attr(d, "value_type");
        if (!names_dry_state(vt, w == 1)) throw H5Error{std::string(nm) + ": value_type is not the dry state: " + vt};
        if (!w) grid = read_str_attr(d, "grid");
        Id t(must(H5Dget_type(d), "H5Dget_type"), H5Tclose);
        if (H5Tget_class(t) != H5T_FLOAT) throw H5Error{std::string(nm) + " is not floating point"};
        int bytes = (int)H5Tget_size(t);
        if (!w) in.float_bytes = bytes;
        else if (bytes != in.float_bytes) throw H5Error{"Y.c and Y.f differ in precision"};
        Id sp(must(H5Dget_space(d), "H5Dget_space"), H5Sclose);
        int nd = H5Sget_simple_extent_ndims(sp);
        if (nd < 1 || nd > 5) throw H5Error{std::string(nm) + ": unexpected rank"};
        std::vector<hsize_t> dims(nd);
        H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
        (w ? fi.fdims : fi.cdims) = dims;
    }
    if (lay[0] != lay[1]) throw H5Error{"Y.c and Y.f have different data layouts"};
    const auto &c = fi.cdims, &f = fi.fdims;
    if (lay[0] == "VF") {
        if (c.size() != 2 || f.size() != 2 || c[0] != 4 || f[0] != 1 || f[1] != c[1] + 1)
            throw H5Error{"VF datasets do not hold (rho, u_h, rho e_tot) and u_3 on nlev + 1 faces"};
        in.kind = CA_IO_COLUMN;
        in.layout = CA_IO_VF;
        in.nlev = (int32_t)c[1];
        in.nq = 1;
        in.nelem = 1;
    } else if (lay[0] == "VIJFH" || lay[0] == "VIJHF") {
        const bool fh = lay[0] == "VIJFH";
        // HDF5 order (h, f, j, i, v) for VIJFH and (f, h, j, i, v) for VIJHF
        const int iF = fh ? 1 : 0, iH = fh ? 0 : 1;
        if (c.size() != 5 || f.size() != 5 || c[iF] != 4 || f[iF] != 1 || c[2] != c[3] || f[2] != c[2] ||
            f[3] != c[3] || f[iH] != c[iH] || f[4] != c[4] + 1)
            throw H5Error{lay[0] + " datasets do not hold (rho, u_h, rho e_tot) and u_3 on nlev + 1 faces"};
        in.kind = CA_IO_SPHERE;
        in.layout = fh ? CA_IO_VIJFH : CA_IO_VIJHF;
        in.nlev = (int32_t)c[4];
        in.nq = (int32_t)c[2];
        in.nelem = (int32_t)c[iH];
        // h_elem from the grid chain, when the file names it
        std::string gp = "grids/" + grid;
        if (H5Lexists(file, "grids", H5P_DEFAULT) > 0 && H5Lexists(file, gp.c_str(), H5P_DEFAULT) > 0) {
            Id gg(must(H5Oopen(file, gp.c_str(), H5P_DEFAULT), "H5Oopen " + gp), H5Oclose);
            if (H5Aexists(gg, "horizontal_grid") > 0) {
                std::string hp = "grids/" + read_str_attr(gg, "horizontal_grid");
                Id hg(must(H5Oopen(file, hp.c_str(), H5P_DEFAULT), "H5Oopen " + hp), H5Oclose);
                std::string tp = "topologies/" + read_str_attr(hg, "topology");
                Id tg(must(H5Oopen(file, tp.c_str(), H5P_DEFAULT), "H5Oopen " + tp), H5Oclose);
                std::string mp = "meshes/" + read_str_attr(tg, "mesh");
                Id mg(must(H5Oopen(file, mp.c_str(), H5P_DEFAULT), "H5Oopen " + mp), H5Oclose);
                if (H5Aexists(mg, "ne") > 0) in.h_elem = (int32_t)read_num_attr(mg, "ne");
            }
        }
    } else {
        throw H5Error{"unsupported data layout " + lay[0]};
    }
}

}  // namespace

extern "C" {

const char *ca_io_last_error(void) { return g_err.c_str(); }

ca_status ca_io_write_state(const char *path, const ca_io_grid *g, const float *c, const float *f, double time,
                            uint64_t atmos_model_hash) {
    if (!path || !g || !c || !f || !g->z_face) return fail(CA_ERR_ARG, "ca_io_write_state: null argument");
    if (g->nlev < 1) return fail(CA_ERR_ARG, "ca_io_write_state: nlev < 1");
    for (int k = 0; k < g->nlev; ++k)
        if (!(g->z_face[k + 1] > g->z_face[k]))
            return fail(CA_ERR_ARG, "ca_io_write_state: z_face not strictly increasing");
    if (g->kind == CA_IO_SPHERE) {
        if (g->h_elem < 1 || g->nq < 2 || !(g->radius > 0))
            return fail(CA_ERR_ARG, "ca_io_write_state: sphere needs h_elem >= 1, nq >= 2 and radius > 0");
        if (g->layout != CA_IO_VIJFH && g->layout != CA_IO_VIJHF)
            return fail(CA_ERR_ARG, "ca_io_write_state: sphere layout must be VIJFH or VIJHF");
        if (g->elemorder) {
            const int64_t nh = 6ll * g->h_elem * g->h_elem;
            for (int64_t e = 0; e < nh; ++e) {
                const int32_t *p = g->elemorder + 3 * e;
                if (p[0] < 1 || p[0] > g->h_elem || p[1] < 1 || p[1] > g->h_elem || p[2] < 1 || p[2] > 6)
                    return fail(CA_ERR_ARG, "ca_io_write_state: elemorder entry out of range");
            }
        }
    } else if (g->kind == CA_IO_COLUMN) {
        if (g->layout != CA_IO_VF) return fail(CA_ERR_ARG, "ca_io_write_state: a column is written as VF");
    } else {
        return fail(CA_ERR_ARG, "ca_io_write_state: unknown grid kind");
    }
    Quiet q;
    try {
        write_file(path, g, c, f, time, atmos_model_hash);
    } catch (const H5Error &e) {
        return fail(CA_ERR_DEVICE, "ca_io_write_state: HDF5: " + e.what);
    }
    return CA_OK;
}

ca_status ca_io_read_info(const char *path, ca_io_info *info) {
    if (!path || !info) return fail(CA_ERR_ARG, "ca_io_read_info: null argument");
    Quiet q;
    try {
        if (H5Fis_hdf5(path) <= 0) return fail(CA_ERR_ARG, std::string("ca_io_read_info: not an HDF5 file: ") + path);
        Id file(must(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), std::string("H5Fopen ") + path), H5Fclose);
        FileInfo fi;
        inspect(file, fi);
        *info = fi.info;
    } catch (const H5Error &e) {
        return fail(CA_ERR_ARG, "ca_io_read_info: " + e.what);
    }
    return CA_OK;
}

ca_status ca_io_read_state(const char *path, float *c, int64_t c_len, float *f, int64_t f_len) {
    if (!path || !c || !f) return fail(CA_ERR_ARG, "ca_io_read_state: null argument");
    Quiet q;
    try {
        if (H5Fis_hdf5(path) <= 0) return fail(CA_ERR_ARG, std::string("ca_io_read_state: not an HDF5 file: ") + path);
        Id file(must(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), std::string("H5Fopen ") + path), H5Fclose);
        FileInfo fi;
        inspect(file, fi);
        const ca_io_info &in = fi.info;
        const int64_t blk_c = (int64_t)in.nlev * in.nq * in.nq, blk_f = (int64_t)(in.nlev + 1) * in.nq * in.nq;
        const int64_t nc = blk_c * 4 * in.nelem, nf = blk_f * in.nelem;
        if (c_len != nc || f_len != nf)
            return fail(CA_ERR_ARG, "ca_io_read_state: buffer sizes " + std::to_string(c_len) + ", " +
                                        std::to_string(f_len) + " differ from the file's " + std::to_string(nc) +
                                        ", " + std::to_string(nf));
        for (int w = 0; w < 2; ++w) {
            const char *nm = w ? "fields/Y.f" : "fields/Y.c";
            Id d(must(H5Dopen2(file, nm, H5P_DEFAULT), std::string("H5Dopen2 ") + nm), H5Dclose);
            float *dst = w ? f : c;
            const int F = w ? 1 : 4;
            const int64_t blk = w ? blk_f : blk_c;
            if (in.layout != CA_IO_VIJHF) {
                must(H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, dst), "H5Dread", nm);
                continue;
            }
            std::vector<float> tmp((size_t)blk * F * in.nelem);
            must(H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()), "H5Dread", nm);
            for (int64_t h = 0; h < in.nelem; ++h)
                for (int qf = 0; qf < F; ++qf)
                    std::memcpy(dst + (h * F + qf) * blk, &tmp[(qf * in.nelem + h) * blk], blk * sizeof(float));
        }
    } catch (const H5Error &e) {
        return fail(CA_ERR_ARG, "ca_io_read_state: " + e.what);
    }
    return CA_OK;
}

}  // extern "C"
