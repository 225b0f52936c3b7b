// VTK XML unstructured-grid writer (include/ensc_vtu.h).
//
// What DataOut::write_vtu (Explicit3D/source/dataOut.cpp:84-195) hands to
// vtkXMLUnstructuredGridWriter, written directly: a VTKFile of type
// UnstructuredGrid, version 1.0, little endian, UInt64 block headers, one
// Piece with the point and cell arrays, the points, the hexahedral cells and
// a TimeValue field array.  Every array is "appended": the XML part gives
// its offset into the raw section that follows, where it is stored as a
// UInt64 byte count and the bytes.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ensc_vtu.h"
#include "guard.h"

namespace ensc {
extern thread_local std::string g_last_error;
}

namespace {

struct Appended {
    std::string xml;           // DataArray elements, in order
    std::vector<char> raw;     // the appended section after '_'
    void add(const std::string& indent, const char* type, const std::string& attrs, const void* data,
             uint64_t bytes) {
        xml += indent + "<DataArray type=\"" + type + "\" " + attrs + " format=\"appended\" offset=\"" +
               std::to_string(raw.size()) + "\"/>\n";
        const char* b = reinterpret_cast<const char*>(&bytes);
        raw.insert(raw.end(), b, b + sizeof(bytes));
        const char* d = static_cast<const char*>(data);
        raw.insert(raw.end(), d, d + bytes);
    }
};

std::string xml_escape(const char* s) {
    std::string o;
    for (; *s; ++s) {
        switch (*s) {
            case '&': o += "&amp;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            case '"': o += "&quot;"; break;
            default: o += *s;
        }
    }
    return o;
}

void add_f32(Appended& A, const std::string& indent, const ensc_vtu_array& a, uint64_t n) {
    std::vector<float> f((size_t)n * a.n_comp);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)a.data[i];
    A.add(indent, "Float32",
          "Name=\"" + xml_escape(a.name) + "\" NumberOfComponents=\"" + std::to_string(a.n_comp) + "\"", f.data(),
          f.size() * sizeof(float));
}

bool check_arrays(const ensc_vtu_array* a, uint32_t n) {
    if (n && !a) return false;
    for (uint32_t i = 0; i < n; ++i)
        if (!a[i].name || !a[i].data || !(a[i].n_comp == 1 || a[i].n_comp == 3 || a[i].n_comp == 6 ||
                                          a[i].n_comp == 9))
            return false;
    return true;
}

}  // namespace

extern "C" ensc_status ensc_vtu_write(const char* path, uint32_t n_points, const double* points, uint32_t n_cells,
                                      const uint32_t* conn, double time, const ensc_vtu_array* point_data,
                                      uint32_t n_point_data, const ensc_vtu_array* cell_data, uint32_t n_cell_data) {
    if (!path || (n_points && !points) || (n_cells && !conn) || !check_arrays(point_data, n_point_data) ||
        !check_arrays(cell_data, n_cell_data)) {
        ensc::g_last_error = "ensc_vtu_write: null or malformed argument (n_comp must be 1, 3, 6 or 9)";
        return ENSC_ERR_ARG;
    }
    return ensc::guarded([&]() -> ensc_status {
        for (uint64_t i = 0; i < (uint64_t)n_cells * 8; ++i)
            if (conn[i] >= n_points) {
                ensc::g_last_error = "ensc_vtu_write: cell " + std::to_string(i / 8) + " names node " +
                                     std::to_string(conn[i]) + " of " + std::to_string(n_points);
                return ENSC_ERR_ARG;
            }
        Appended A;
        const std::string i4 = "        ", i5 = "          ";
        std::string field, pdata, cdata, pts, cells;
        {
            const float t = (float)time;
            A.add(i4, "Float32", "Name=\"TimeValue\" NumberOfTuples=\"1\"", &t, sizeof(t));
            field.swap(A.xml);
        }
        for (uint32_t i = 0; i < n_point_data; ++i) add_f32(A, i5, point_data[i], n_points);
        pdata.swap(A.xml);
        for (uint32_t i = 0; i < n_cell_data; ++i) add_f32(A, i5, cell_data[i], n_cells);
        cdata.swap(A.xml);
        {
            const ensc_vtu_array p{"Points", 3, 0, points};
            std::vector<float> f((size_t)n_points * 3);
            for (size_t i = 0; i < f.size(); ++i) f[i] = (float)p.data[i];
            A.add(i5, "Float32", "Name=\"Points\" NumberOfComponents=\"3\"", f.data(), f.size() * sizeof(float));
            pts.swap(A.xml);
        }
        {
            std::vector<int64_t> c((size_t)n_cells * 8), off(n_cells);
            for (size_t i = 0; i < c.size(); ++i) c[i] = conn[i];
            for (size_t e = 0; e < n_cells; ++e) off[e] = (int64_t)(8 * (e + 1));
            std::vector<uint8_t> ty(n_cells, 12);  // VTK_HEXAHEDRON
            A.add(i5, "Int64", "Name=\"connectivity\"", c.data(), c.size() * sizeof(int64_t));
            A.add(i5, "Int64", "Name=\"offsets\"", off.data(), off.size() * sizeof(int64_t));
            A.add(i5, "UInt8", "Name=\"types\"", ty.data(), ty.size());
            cells.swap(A.xml);
        }
        std::string head = "<?xml version=\"1.0\"?>\n"
                           "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                           "header_type=\"UInt64\">\n"
                           "  <UnstructuredGrid>\n"
                           "    <FieldData>\n" +
                           field + "    </FieldData>\n" + "    <Piece NumberOfPoints=\"" + std::to_string(n_points) +
                           "\" NumberOfCells=\"" + std::to_string(n_cells) + "\">\n" + "      <PointData>\n" + pdata +
                           "      </PointData>\n" + "      <CellData>\n" + cdata + "      </CellData>\n" +
                           "      <Points>\n" + pts + "      </Points>\n" + "      <Cells>\n" + cells +
                           "      </Cells>\n" + "    </Piece>\n" + "  </UnstructuredGrid>\n" +
                           "  <AppendedData encoding=\"raw\">\n   _";
        const std::string tail = "\n  </AppendedData>\n</VTKFile>\n";
        FILE* f = std::fopen(path, "wb");
        if (!f) {
            ensc::g_last_error = std::string("ensc_vtu_write: cannot open ") + path;
            return ENSC_ERR_ARG;
        }
        bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
        ok = ok && std::fwrite(A.raw.data(), 1, A.raw.size(), f) == A.raw.size();
        ok = ok && std::fwrite(tail.data(), 1, tail.size(), f) == tail.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) {
            ensc::g_last_error = std::string("ensc_vtu_write: write failed for ") + path;
            return ENSC_ERR_ARG;
        }
        return ENSC_OK;
    });
}
