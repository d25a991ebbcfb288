/*
 * caio.h -- ClimaCore HDF5 checkpoint I/O for the native dry-dycore state
 * (host side, libcaio.so, linked against libhdf5).
 *
 * Replaces the two ends of the reference's checkpoint / restart path:
 *
 *   ca_io_write_state   save_state_to_disk_func(integrator, output_dir)
 *                       src/callbacks/callbacks.jl:297-320: an HDF5 file with
 *                       the root attributes "time" (FT) and "atmos_model_hash"
 *                       (UInt64) and the FieldVector Y written by ClimaCore's
 *                       InputOutput.write!(HDF5Writer, Y, "Y")
 *   ca_io_read_info     get_state_restart(restart_file, ...)
 *   ca_io_read_state      src/simulation/restart.jl:22-43: InputOutput.read_field(reader, "Y"),
 *                       read_attribute(file, "time") and the model-hash check
 *
 * The file layout is ClimaCore 0.15.1's InputOutput format (the version the
 * reference pins in .buildkite/Manifest-v1.11.toml): "fields/Y" a group with
 * attributes type = "FieldVector" and fields = ["c", "f"]; "fields/Y.c" and
 * "fields/Y.f" datasets holding the parent arrays of Y.c and Y.f (Julia
 * column-major, so the HDF5 dimensions are the Julia ones reversed) with
 * attributes type = "Field", data_layout, value_type, grid and staggering;
 * the grid under "grids/", "topologies/", "meshes/" and "domains/".
 *
 * Arrays cross the ABI in host memory, in ClimaCore's parent-array order:
 *   sphere (CA_IO_SPHERE): VIJFH, Y.c as Julia (Nv, Nq, Nq, 4, Nh) and Y.f as
 *       (Nv + 1, Nq, Nq, 1, Nh); element h is the h-th element of the file
 *       (the topology's element order);
 *   column (CA_IO_COLUMN): VF, Y.c as Julia (Nv, 4) and Y.f as (Nv + 1, 1).
 * ca_unpack_state / ca_pack_state (cadycore.h) move them to and from the
 * native device layout.  Centre components are (rho, u_1, u_2, rho e_tot)
 * with u_h covariant (C12); the face component is u_3 (C3).
 *
 * Every call returns a ca_status; ca_io_last_error() describes the last
 * failure of the calling thread.  Nothing throws across the ABI.
 */
#ifndef CAIO_H
#define CAIO_H
#include <stdint.h>

#include "cadycore.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ca_io_kind { CA_IO_SPHERE = 0, CA_IO_COLUMN = 1 } ca_io_kind;
typedef enum ca_io_layout { CA_IO_VIJFH = 0, CA_IO_VIJHF = 1, CA_IO_VF = 2 } ca_io_layout;

typedef struct ca_io_grid {
    int32_t kind;              /* ca_io_kind                                          */
    int32_t layout;            /* sphere: CA_IO_VIJFH or CA_IO_VIJHF for the file's
                                  datasets (arrays still cross the ABI as VIJFH);
                                  column: CA_IO_VF                                     */
    int32_t nlev;              /* cell centres per column                             */
    int32_t h_elem;            /* sphere: elements along a cube edge (Nh = 6 h_elem^2) */
    int32_t nq;                /* sphere: GLL nodes along an element edge              */
    int32_t deep;              /* sphere: deep-atmosphere metric (DeepSphericalGlobalGeometry) */
    double radius;             /* sphere: planet radius [m]                            */
    const double *z_face;      /* nlev + 1 face heights [m], increasing                */
    const int32_t *elemorder;  /* sphere: 3 x Nh (x, y, panel; 1-based) of each element
                                  in file order, or NULL for the panel-major order
                                  h = (panel - 1) h_elem^2 + (y - 1) h_elem + x - 1    */
} ca_io_grid;

typedef struct ca_io_info {
    int32_t kind;              /* ca_io_kind                                           */
    int32_t layout;            /* ca_io_layout of the file's datasets                  */
    int32_t nlev;              /* cell centres                                         */
    int32_t nq;                /* nodes along an element edge (1 for a column)          */
    int32_t nelem;             /* elements (1 for a column)                             */
    int32_t h_elem;            /* from the mesh group; 0 when absent                    */
    int32_t float_bytes;       /* 4 or 8: the element size of Y in the file             */
    int32_t has_hash;          /* the root attribute "atmos_model_hash" is present      */
    double time;               /* root attribute "time" [s]                             */
    uint64_t atmos_model_hash;
} ca_io_info;

/* Writes Y (host arrays as described above), time (stored as Float32, the
 * reference's FT) and atmos_model_hash to a new file at path (an existing
 * file is replaced).  CA_ERR_ARG on bad sizes or pointers, CA_ERR_DEVICE on
 * an HDF5 failure (message from the HDF5 call that failed). */
ca_status ca_io_write_state(const char *path, const ca_io_grid *grid, const float *c, const float *f,
                            double time, uint64_t atmos_model_hash);
/* Reads the sizes, layout, time and hash of a checkpoint.  CA_ERR_ARG when
 * the file is not a dry-dycore FieldVector checkpoint (another value type,
 * a missing dataset or attribute). */
ca_status ca_io_read_info(const char *path, ca_io_info *info);
/* Reads Y.c and Y.f into host arrays of c_len and f_len floats, converting
 * Float64 files to float32 and VIJHF datasets to VIJFH order. */
ca_status ca_io_read_state(const char *path, float *c, int64_t c_len, float *f, int64_t f_len);
const char *ca_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
