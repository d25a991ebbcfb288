/* A second, independent writer of a dry-state checkpoint, used by
 * tests/test_restart_io.py to read files that libcaio did not write.  It
 * writes what a Float64 ClimaAtmos run's file holds in the places where it
 * differs from libcaio's own output: Float64 datasets and time, the VIJHF
 * data layout (Julia (Nv, Ni, Nj, Nh, F)), variable-length string attributes,
 * and no "atmos_model_hash".  Values: Y.c element (v, i, j, h, f) (0-based)
 * = 1000 f + 100 h + 10 j + i + v / 100, Y.f = -(the same with f = 0).
 *
 * usage: h5_foreign_writer out.hdf5 h_elem nlev [value_type_of_Y.c]
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void vattr(hid_t obj, const char *name, const char *v) {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, H5T_VARIABLE);
    H5Tset_cset(t, H5T_CSET_UTF8);
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, t, &v);
    H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
}

static void field(hid_t file, hid_t lcpl, const char *path, int F, int nh, int nv, double sign, const char *vt,
                  const char *stag) {
    hsize_t dims[5] = {(hsize_t)F, (hsize_t)nh, 4, 4, (hsize_t)nv}; /* HDF5 order of Julia (Nv, Ni, Nj, Nh, F) */
    size_t n = (size_t)F * nh * 16 * nv;
    double *buf = malloc(n * sizeof(double));
    size_t p = 0;
    for (int f = 0; f < F; ++f)
        for (int h = 0; h < nh; ++h)
            for (int j = 0; j < 4; ++j)
                for (int i = 0; i < 4; ++i)
                    for (int v = 0; v < nv; ++v)
                        buf[p++] = sign * (1000.0 * f + 100.0 * h + 10.0 * j + i + v / 100.0);
    hid_t sp = H5Screate_simple(5, dims, NULL);
    hid_t d = H5Dcreate2(file, path, H5T_IEEE_F64LE, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    vattr(d, "type", "Field");
    vattr(d, "data_layout", "VIJHF");
    vattr(d, "value_type", vt);
    vattr(d, "grid", "center_extruded_finite_difference_grid");
    vattr(d, "staggering", stag);
    H5Dclose(d);
    H5Sclose(sp);
    free(buf);
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    int he = atoi(argv[2]), nv = atoi(argv[3]), nh = 6 * he * he;
    const char *vtc = argc > 4 ? argv[4]
                               : "@NamedTuple{\xCF\x81::Float64, u\xE2\x82\x95::ClimaCore.Geometry.Covariant12Vector{Float64}, "
                                 "\xCF\x81" "e_tot::Float64}";
    /* the @NamedTuple form names fields as "ρ::", which the reader must accept too */
    hid_t file = H5Fcreate(argv[1], H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return 1;
    hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    double t = 1234.5;
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(file, "time", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_DOUBLE, &t);
    H5Aclose(a);
    H5Sclose(sp);
    hid_t g = H5Gcreate2(file, "fields/Y", lcpl, H5P_DEFAULT, H5P_DEFAULT);
    vattr(g, "type", "FieldVector");
    {
        const char *names[2] = {"c", "f"};
        hid_t t = H5Tcopy(H5T_C_S1);
        H5Tset_size(t, H5T_VARIABLE);
        hsize_t two = 2;
        hid_t s = H5Screate_simple(1, &two, NULL);
        hid_t at = H5Acreate2(g, "fields", t, s, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(at, t, names);
        H5Aclose(at);
        H5Sclose(s);
        H5Tclose(t);
    }
    H5Gclose(g);
    field(file, lcpl, "fields/Y.c", 4, nh, nv, 1.0, vtc, "CellCenter");
    field(file, lcpl, "fields/Y.f", 1, nh, nv + 1, -1.0,
          "@NamedTuple{u\xE2\x82\x83::ClimaCore.Geometry.Covariant3Vector{Float64}}", "CellFace");
    H5Pclose(lcpl);
    return H5Fclose(file) < 0;
}
