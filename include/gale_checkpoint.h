/*
 * gale_checkpoint.h -- reading the reference's HDF5 checkpoints (FeVariable
 * fields, material swarms, time information) without an HDF5 library.
 * Exported by libgale_stokes.so next to gale_stokes.h.
 *
 * The reader covers the part of the HDF5 file format that the reference's
 * writers produce: superblock version 0 or 1 with 8-byte offsets and lengths,
 * version-1 object headers, a symbol-table root group, and contiguous or
 * compact datasets of signed 32/64-bit integers or 32/64-bit floats in either
 * byte order.  Other files (chunked or filtered datasets, version-2 object
 * headers, other types) return GS_H5_UNSUPPORTED.  Each call opens, reads (or
 * writes) and closes the file; nothing is kept between calls.
 */
#ifndef GALE_CHECKPOINT_H
#define GALE_CHECKPOINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GS_H5_I32 = 1, GS_H5_I64 = 2, GS_H5_F32 = 3, GS_H5_F64 = 4 };
enum {
  GS_H5_IO = -1,          /* file cannot be opened or read */
  GS_H5_FORMAT = -2,      /* not an HDF5 file, or a structure points outside the file */
  GS_H5_UNSUPPORTED = -3, /* HDF5, but outside the subset above */
  GS_H5_NOT_FOUND = -4,   /* no such dataset / attribute in the root group */
  GS_H5_SHAPE = -5,       /* dataset shape does not match the request */
  GS_H5_INVALID = -6      /* bad argument (null pointer, capacity too small) */
};
#define GS_H5_MAX_RANK 4

/* Rank, dimensions (dims[GS_H5_MAX_RANK], unused ones 0) and element type of
 * root-group dataset `name` (H5Dopen + H5Dget_space + H5Dget_type).  Returns 0
 * or a GS_H5_* error. */
int32_t gs_h5_dataset_info(const char* path, const char* name, int32_t* rank, int64_t* dims, int32_t* type);

/* All elements of dataset `name`, converted to double / int32 (H5Dread with
 * H5T_NATIVE_DOUBLE / H5T_NATIVE_INT), row-major.  Returns the element count
 * or a GS_H5_* error (GS_H5_INVALID if capacity is smaller than the count). */
int64_t gs_h5_read_f64(const char* path, const char* name, double* out, int64_t capacity);
int64_t gs_h5_read_i32(const char* path, const char* name, int32_t* out, int64_t capacity);

/* Integer attribute `name` of the root group (H5Aopen + H5Aread), up to
 * capacity values.  Returns its element count or a GS_H5_* error. */
int32_t gs_h5_attribute_i64(const char* path, const char* name, int64_t* out, int32_t capacity);

/* FeVariable_ReadFromFile (StgFEM/Discretisation/src/FeVariable.cxx:2186-2390):
 * the `dofs` components of a field at n_nodes nodes (global node order, x
 * fastest) from "/data".  Both checkpoint layouts: version 1 (no "checkpoint
 * file version" attribute) has the node index in the first column, version 2
 * has none; columns beyond the components are the saved vertex coordinates and
 * are skipped.  A P0 pressure field is read with dofs 1 and n_nodes = elements.
 * out: n_nodes*dofs.  Returns n_nodes or a GS_H5_* error. */
int64_t gs_h5_read_field(const char* path, int32_t dim, int32_t dofs, int64_t n_nodes, double* out);

/* FileParticleLayout (StgDomain/Swarm/src/FileParticleLayout.cxx:333-470)
 * reading a SwarmDump file (StgDomain/Swarm/src/SwarmDump.cxx:213-330): the
 * positions (P x 3, the third 0 in 2D) and material indices of the swarm's
 * particles, from "/Position" and "/MaterialIndex", or the older
 * "/<swarm>-Position" and "/<swarm>-MaterialIndex".  With coord and material
 * both NULL returns P only.  Returns P or a GS_H5_* error. */
int64_t gs_h5_read_swarm(const char* path, const char* swarm, int32_t dim, int64_t capacity, double* coord,
                         int32_t* material);

/* Writers of the same subset: one symbol-table root group of contiguous
 * datasets plus root attributes (at most 8 datasets per file).  Returns 0 or
 * a GS_H5_* error.  The files are read back by the functions above. */

/* FeVariable_SaveToFile (StgFEM/Discretisation/src/FeVariable.cxx:1938-2130),
 * checkpoint version 2: "/data" of n_nodes rows, each the node's coordinates
 * (coords P*3 as elsewhere in this library, or NULL for none) then its `dofs`
 * components (values n_nodes*dofs); root attributes "checkpoint file
 * version" = 2, "dimensions" and, if mesh_res is given, "mesh resolution". */
int32_t gs_h5_write_field(const char* path, int32_t dim, int32_t dofs, int64_t n_nodes, const double* values,
                          const double* coords, const int32_t* mesh_res);

/* SwarmDump (StgDomain/Swarm/src/SwarmDump.cxx:213-330) of a material swarm:
 * "/Position" (P x dim doubles from coord, P*3), "/MaterialIndex" (P x 1
 * int32) and the root attribute "Swarm Particle Count". */
int32_t gs_h5_write_swarm(const char* path, int32_t dim, int64_t P, const double* coord, const int32_t* material);
/* The same with n_vars (at most 6) further scalar swarm variables, each a
 * "/<name>" dataset of P doubles (e.g. "PostFailureWeakening", the plastic
 * strain of StrainWeakening). */
int32_t gs_h5_write_swarm_vars(const char* path, int32_t dim, int64_t P, const double* coord, const int32_t* material,
                               int32_t n_vars, const char* const* names, const double* const* values);
/* One scalar swarm variable (P values) by name, as FileParticleLayout looks it
 * up: "/<name>", then "/<swarm>-<name>".  out NULL returns P only.  Returns P
 * or a GS_H5_* error. */
int64_t gs_h5_read_swarm_var(const char* path, const char* swarm, const char* name, int64_t capacity, double* out);

/* The timeInfo file (read by FileParticleLayout.cxx:565-597 on restart):
 * "/Dt", "/currentTime" (doubles) and "/nproc" (int32). */
int32_t gs_h5_write_time(const char* path, double dt, double time, int32_t nproc);

#ifdef __cplusplus
}
#endif
#endif
