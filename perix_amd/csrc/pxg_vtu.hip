// pxg_vtu.hip -- binary VTU output of PD results (host code).
//
// The content and order of PDProblem::savePDResults (SaveResults.cpp:53-204):
// N points and N VTK_VERTEX cells; point data ElmtID (the parent cell id,
// -1 for ghosts), one Float64 scalar per dof named by the DOFs block, then
// the requested projections (1, 3 or 9 components); an empty CellData block;
// the point coordinates; connectivity 0..N-1, offsets 1..N, types 1.  The
// reference writes every array as ASCII text with 16 significant digits; this
// writer appends them as raw little-endian binary (UInt64 byte-count headers),
// which keeps every value exactly and is read by any VTK XML reader.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include "pxg_internal.h"

namespace {

struct Block {
    const void *data;
    uint64_t bytes;
    std::vector<char> owned;   // for arrays built here
};

std::string xml_escape(const char *s)
{
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

} // namespace

extern "C" int pxg_write_vtu(const char *path, int n, int nbulk, const double *coords, int ndof,
                             const char *const *dof_names, const double *u, int nfields,
                             const char *const *field_names, const int *field_comps,
                             const double *const *field_data)
{
    PXG_CHECK(path && n >= 0 && nbulk >= 0 && nbulk <= n && coords && ndof >= 0 && nfields >= 0,
              PXG_ERR_ARG, "pxg_write_vtu: bad arguments");
    PXG_CHECK(ndof == 0 || (dof_names && u), PXG_ERR_ARG, "pxg_write_vtu: dof names and values required");
    PXG_CHECK(nfields == 0 || (field_names && field_comps && field_data), PXG_ERR_ARG,
              "pxg_write_vtu: field names, components and values required");
    for (int f = 0; f < nfields; f++)
        PXG_CHECK(field_names[f] && field_data[f] && (field_comps[f] == 1 || field_comps[f] == 3 || field_comps[f] == 9),
                  PXG_ERR_ARG, "pxg_write_vtu: a field needs a name, values and 1, 3 or 9 components");
    for (int a = 0; a < ndof; a++) PXG_CHECK(dof_names[a], PXG_ERR_ARG, "pxg_write_vtu: missing dof name");
    const size_t N = (size_t)n;
    std::vector<Block> blocks;
    std::string xml;
    uint64_t off = 0;
    auto add = [&](const std::string &head, Block b) {
        xml += head + " format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
        off += sizeof(uint64_t) + b.bytes;
        blocks.push_back(std::move(b));
    };
    auto owned = [](size_t bytes) {
        Block b;
        b.owned.resize(bytes);
        b.data = nullptr;
        b.bytes = bytes;
        return b;
    };
    xml += "<?xml version=\"1.0\"?>\n";
    xml += "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    xml += "  <UnstructuredGrid>\n";
    xml += "    <Piece NumberOfPoints=\"" + std::to_string(N) + "\" NumberOfCells=\"" + std::to_string(N) + "\">\n";
    xml += "      <PointData Scalars=\"" + (ndof > 0 ? xml_escape(dof_names[0]) : std::string("ElmtID")) + "\">\n";
    {
        Block b = owned(N * sizeof(int64_t));
        int64_t *id = reinterpret_cast<int64_t *>(b.owned.data());
        for (size_t i = 0; i < N; i++) id[i] = (i < (size_t)nbulk) ? (int64_t)(i + 1) : -1;
        add("        <DataArray type=\"Int64\" Name=\"ElmtID\" NumberOfComponents=\"1\"", std::move(b));
    }
    for (int a = 0; a < ndof; a++) {
        Block b = owned(N * sizeof(double));
        double *v = reinterpret_cast<double *>(b.owned.data());
        for (size_t i = 0; i < N; i++) v[i] = u[i * (size_t)ndof + a];
        add("        <DataArray type=\"Float64\" Name=\"" + xml_escape(dof_names[a]) + "\" NumberOfComponents=\"1\"",
            std::move(b));
    }
    for (int f = 0; f < nfields; f++) {
        Block b;
        b.data = field_data[f];
        b.bytes = N * (size_t)field_comps[f] * sizeof(double);
        add("        <DataArray type=\"Float64\" Name=\"" + xml_escape(field_names[f]) +
                "\" NumberOfComponents=\"" + std::to_string(field_comps[f]) + "\"",
            std::move(b));
    }
    xml += "      </PointData>\n";
    xml += "      <CellData>\n      </CellData>\n";
    xml += "      <Points>\n";
    {
        Block b;
        b.data = coords;
        b.bytes = N * 3 * sizeof(double);
        add("        <DataArray type=\"Float64\" NumberOfComponents=\"3\"", std::move(b));
    }
    xml += "      </Points>\n";
    xml += "      <Cells>\n";
    {
        Block c = owned(N * sizeof(int64_t)), o = owned(N * sizeof(int64_t)), t = owned(N);
        int64_t *cc = reinterpret_cast<int64_t *>(c.owned.data());
        int64_t *oo = reinterpret_cast<int64_t *>(o.owned.data());
        for (size_t i = 0; i < N; i++) { cc[i] = (int64_t)i; oo[i] = (int64_t)(i + 1); }
        for (size_t i = 0; i < N; i++) t.owned[i] = 1;   // VTK_VERTEX
        add("        <DataArray type=\"Int64\" Name=\"connectivity\"", std::move(c));
        add("        <DataArray type=\"Int64\" Name=\"offsets\"", std::move(o));
        add("        <DataArray type=\"UInt8\" Name=\"types\"", std::move(t));
    }
    xml += "      </Cells>\n";
    xml += "    </Piece>\n";
    xml += "  </UnstructuredGrid>\n";
    xml += "  <AppendedData encoding=\"raw\">\n   _";
    FILE *fp = std::fopen(path, "wb");
    PXG_CHECK(fp, PXG_ERR_ARG, "pxg_write_vtu: cannot open the output file");
    bool ok = std::fwrite(xml.data(), 1, xml.size(), fp) == xml.size();
    for (const Block &b : blocks) {
        const uint64_t h = b.bytes;
        const void *d = b.owned.empty() ? b.data : b.owned.data();
        ok = ok && std::fwrite(&h, sizeof h, 1, fp) == 1;
        ok = ok && (b.bytes == 0 || std::fwrite(d, 1, b.bytes, fp) == b.bytes);
    }
    const std::string tail = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail.data(), 1, tail.size(), fp) == tail.size();
    ok = (std::fclose(fp) == 0) && ok;
    PXG_CHECK(ok, PXG_ERR_ARG, "pxg_write_vtu: write failed");
    return PXG_OK;
}
