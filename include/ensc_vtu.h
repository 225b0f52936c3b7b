/*
 * ensc_vtu.h - VTK XML unstructured-grid (.vtu) writer for hexahedral
 * meshes, without VTK.
 *
 * Replaces DataOut::write_vtu (Explicit3D/source/dataOut.cpp:84-195,
 * Explicit3D/include/dataOut.h:44-97), which builds a vtkUnstructuredGrid
 * and writes it with vtkXMLUnstructuredGridWriter.  The file holds what that
 * writer stores for the reference's grid: the points as Float32 triples,
 * one VTK_HEXAHEDRON (type 12) per element with its 8 node ids (Int64
 * connectivity and offsets, UInt8 types), every point and cell array as
 * Float32 with its name and component count, and a "TimeValue" Float32
 * field-data array.  Arrays are stored uncompressed in the appended raw
 * section (the reference asks VTK for zlib blocks; VTK readers accept
 * both).
 */
#ifndef ENSC_VTU_H
#define ENSC_VTU_H

#include <stdint.h>
#include "ensc_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One named array: n_comp values per point or cell, [n][n_comp] doubles
 * (written as Float32).  n_comp 1: scalar, 3: vector, 6: symmetric tensor
 * in the component order the caller gives (the reference passes
 * xx yy zz xy yz xz). */
typedef struct ensc_vtu_array {
    const char* name;
    int32_t n_comp;
    int32_t pad_;
    const double* data;
} ensc_vtu_array;

/* Write `path`.  points [n_points][3]; conn [n_cells][8] (0-based node ids in
 * Abaqus C3D8R order, which is VTK's hexahedron order); time goes to the
 * TimeValue field array. */
ensc_status ensc_vtu_write(const char* path, uint32_t n_points, const double* points, uint32_t n_cells,
                           const uint32_t* conn, double time, const ensc_vtu_array* point_data,
                           uint32_t n_point_data, const ensc_vtu_array* cell_data, uint32_t n_cell_data);

#ifdef __cplusplus
}
#endif

#endif /* ENSC_VTU_H */
