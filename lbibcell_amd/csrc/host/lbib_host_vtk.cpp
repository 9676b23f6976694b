// lbib_host_vtk.cpp -- VTK XML files (ASCII) without the VTK library: the
// geometry loader Geometry::loadGeometryVTK (Geometry.cpp:162-306) and the
// multiblock writers of vtkFluidReporter (vtkFluidReporter.cpp:44-108),
// vtkCDEReporter (vtkCDEReporter.cpp:44-100) and vtkCellReporter
// (vtkCellReporter.cpp:42-150).
//
// Files follow the layout vtkXMLMultiBlockDataWriter produces: <name>.vtm
// lists DataSet blocks stored as <name>/<base>_<index>.vti|.vtp.  Every
// reporter appends its block to the .vtm of the time step when the file
// already exists, as the reference's read-then-write does.  Values are
// written with 17 significant digits (lossless for FP64).
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "lbib_host.hpp"

namespace LbmLib {
namespace vtkxml {

namespace {

std::string read_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) LBIB_FAIL("Cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

std::string dirname_of(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : path.substr(0, k);
}
std::string basename_of(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? path : path.substr(k + 1);
}

// value of attribute `key` inside one tag's text ("" if absent)
std::string attr(const std::string& tag, const std::string& key) {
    const std::string pat = " " + key + "=\"";
    const size_t a = tag.find(pat);
    if (a == std::string::npos) return "";
    const size_t b = tag.find('"', a + pat.size());
    return tag.substr(a + pat.size(), b - a - pat.size());
}

struct Array {
    std::string name, type;
    int ncomp = 1;
    std::vector<double> num;
    std::vector<std::string> str;   // String arrays: character codes, 0-terminated
};

// the DataArray / Array elements between [from, to) of doc
std::vector<Array> arrays_in(const std::string& doc, size_t from, size_t to) {
    std::vector<Array> out;
    size_t p = from;
    while (true) {
        size_t a = doc.find("<DataArray", p), b = doc.find("<Array", p);
        size_t s = std::min(a, b);
        if (s == std::string::npos || s >= to) break;
        const bool isData = s == a;
        const size_t tagEnd = doc.find('>', s);
        const std::string tag = doc.substr(s, tagEnd - s + 1);
        const std::string closeTag = isData ? "</DataArray>" : "</Array>";
        const bool selfClosed = doc[tagEnd - 1] == '/';
        const size_t bodyEnd = selfClosed ? tagEnd : doc.find(closeTag, tagEnd);
        Array arr;
        arr.name = attr(tag, "Name");
        arr.type = attr(tag, "type");
        const std::string nc = attr(tag, "NumberOfComponents");
        if (!nc.empty()) arr.ncomp = std::atoi(nc.c_str());
        const std::string fmt = attr(tag, "format");
        if (!fmt.empty() && fmt != "ascii")
            LBIB_FAIL("only ASCII VTK XML arrays are supported (array " + arr.name + " is " + fmt + ")");
        if (!selfClosed) {
            const char* c = doc.c_str() + tagEnd + 1;
            const char* e = doc.c_str() + bodyEnd;
            const bool f32 = arr.type == "Float32";
            while (c < e) {
                char* nx = nullptr;
                errno = 0;
                const double v = std::strtod(c, &nx);
                if (nx == c) { c++; continue; }
                // Float32 arrays hold float values (the reference reads them as float)
                arr.num.push_back(f32 ? (double)(float)v : v);
                c = nx;
            }
            if (arr.type == "String") {
                std::string cur;
                for (double v : arr.num) {
                    if (v == 0.0) { arr.str.push_back(cur); cur.clear(); }
                    else cur.push_back((char)(int)v);
                }
                arr.num.clear();
            }
        }
        out.push_back(std::move(arr));
        p = selfClosed ? tagEnd + 1 : bodyEnd + closeTag.size();
    }
    return out;
}

// [begin, end) of the first <section> ... </section> at or after `from`
bool section(const std::string& doc, const std::string& name, size_t from, size_t& b, size_t& e) {
    b = doc.find("<" + name, from);
    if (b == std::string::npos) return false;
    e = doc.find("</" + name + ">", b);
    if (e == std::string::npos) {   // <Name/> empty element
        e = doc.find('>', b);
    }
    return true;
}

struct PolyData {
    long npoints = 0, npolys = 0;
    std::vector<Array> pointData;
    std::vector<double> points;
    std::vector<long> conn, offs;
    const Array* find(const std::string& name) const {
        for (auto& a : pointData)
            if (a.name == name) return &a;
        return nullptr;
    }
};

PolyData read_vtp(const std::string& path) {
    const std::string doc = read_file(path);
    if (doc.find("type=\"PolyData\"") == std::string::npos) LBIB_FAIL(path + " is not a PolyData file");
    PolyData pd;
    size_t b, e;
    if (!section(doc, "Piece", 0, b, e)) LBIB_FAIL(path + ": no Piece");
    const std::string piece = doc.substr(b, doc.find('>', b) - b + 1);
    pd.npoints = std::atol(attr(piece, "NumberOfPoints").c_str());
    pd.npolys = std::atol(attr(piece, "NumberOfPolys").c_str());
    if (section(doc, "PointData", b, b, e)) pd.pointData = arrays_in(doc, b, e);
    if (section(doc, "Points", 0, b, e)) {
        auto a = arrays_in(doc, b, e);
        if (!a.empty()) pd.points = a[0].num;
    }
    if (pd.npolys > 0 && section(doc, "Polys", 0, b, e)) {
        for (auto& a : arrays_in(doc, b, e)) {
            if (a.name == "connectivity") for (double v : a.num) pd.conn.push_back((long)v);
            if (a.name == "offsets") for (double v : a.num) pd.offs.push_back((long)v);
        }
    }
    if ((long)pd.points.size() != 3 * pd.npoints) LBIB_FAIL(path + ": point count does not match NumberOfPoints");
    if ((long)pd.offs.size() != pd.npolys) LBIB_FAIL(path + ": polygon offsets do not match NumberOfPolys");
    return pd;
}

// DataSet file names listed by a .vtm, resolved against its directory
std::vector<std::string> vtm_blocks(const std::string& path) {
    const std::string doc = read_file(path);
    std::vector<std::string> files;
    size_t p = 0;
    while ((p = doc.find("<DataSet", p)) != std::string::npos) {
        const size_t e = doc.find('>', p);
        const std::string f = attr(doc.substr(p, e - p + 1), "file");
        if (!f.empty()) files.push_back(dirname_of(path) + "/" + f);
        p = e;
    }
    return files;
}

void make_dir(const std::string& d) {
    struct stat st;
    if (stat(d.c_str(), &st) == 0) return;
    if (mkdir(d.c_str(), 0755) != 0) LBIB_FAIL("cannot create directory " + d);
}

void put(std::ostream& o, double v) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", v);
    o << b;
}

void write_numbers(std::ostream& o, const std::vector<double>& v) {
    for (size_t k = 0; k < v.size(); k++) {
        o << (k % 6 == 0 ? "\n          " : " ");
        put(o, v[k]);
    }
    o << "\n";
}

void write_array(std::ostream& o, const std::string& type, const std::string& name, int ncomp,
                 const std::vector<double>& v) {
    o << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\"";
    if (ncomp > 1) o << " NumberOfComponents=\"" << ncomp << "\"";
    o << " format=\"ascii\">";
    write_numbers(o, v);
    o << "        </DataArray>\n";
}

// Adds `blockFile` (relative to the .vtm's directory) to <name>.vtm,
// keeping the blocks already listed there.
void append_block(const std::string& vtm, const std::string& blockFile) {
    std::vector<std::string> rel;
    struct stat st;
    if (stat(vtm.c_str(), &st) == 0) {
        const std::string dir = dirname_of(vtm) + "/";
        for (auto& f : vtm_blocks(vtm)) rel.push_back(f.substr(dir.size()));
    }
    rel.push_back(blockFile);
    std::ofstream o(vtm);
    if (!o.is_open()) LBIB_FAIL("cannot write " + vtm);
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "  <vtkMultiBlockDataSet>\n";
    for (size_t k = 0; k < rel.size(); k++)
        o << "    <DataSet index=\"" << k << "\" file=\"" << rel[k] << "\">\n    </DataSet>\n";
    o << "  </vtkMultiBlockDataSet>\n</VTKFile>\n";
}

// next block file of the time step's .vtm: <prefix>_<time>/<base>_<time>_<k>.<ext>
std::string next_block(const std::string& prefix, unsigned time, const std::string& ext, std::string& vtm) {
    vtm = prefix + "_" + std::to_string(time) + ".vtm";
    size_t k = 0;
    struct stat st;
    if (stat(vtm.c_str(), &st) == 0) k = vtm_blocks(vtm).size();
    const std::string base = basename_of(prefix) + "_" + std::to_string(time);
    make_dir(prefix + "_" + std::to_string(time));
    return base + "/" + base + "_" + std::to_string(k) + "." + ext;
}

}  // namespace

void write_image(const std::string& prefix, unsigned time, int nx, int ny, int coarse,
                 const std::vector<std::pair<std::string, std::pair<int, std::vector<double>>>>& fields) {
    std::string vtm;
    const std::string rel = next_block(prefix, time, "vti", vtm);
    const int cx = nx / coarse, cy = ny / coarse;
    std::ofstream o(dirname_of(vtm) + "/" + rel);
    if (!o.is_open()) LBIB_FAIL("cannot write " + rel);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "  <ImageData WholeExtent=\"0 " << cx - 1 << " 0 " << cy - 1 << " 0 0\" Origin=\"0 0 0\" Spacing=\""
      << coarse << " " << coarse << " " << coarse << "\">\n"
      << "  <Piece Extent=\"0 " << cx - 1 << " 0 " << cy - 1 << " 0 0\">\n    <PointData>\n";
    for (auto& f : fields) write_array(o, "Float64", f.first, f.second.first, f.second.second);
    o << "    </PointData>\n    <CellData>\n    </CellData>\n  </Piece>\n  </ImageData>\n</VTKFile>\n";
    o.close();
    append_block(vtm, rel);
}

void write_cells(const std::string& prefix, unsigned time, const geometry::Geometry& geo,
                 const std::map<unsigned, unsigned>& cellTypes) {
    write_cells(prefix, time, geo.getConnections(), cellTypes);
}

void write_cells(const std::string& prefix, unsigned time,
                 const std::vector<std::shared_ptr<geometry::Connection>>& connections,
                 const std::map<unsigned, unsigned>& cellTypes) {
    std::string vtm;
    const std::string rel = next_block(prefix, time, "vtp", vtm);
    // vtkCellReporter.cpp:77-120: one polygon per domain, walked from the
    // domain's first connection (Geometry::connections_ order)
    std::map<unsigned, const geometry::Connection*> first;
    for (auto& c : connections)
        if (!first.count(c->domain)) first[c->domain] = c.get();
    std::vector<double> pts, dom, ct, desc;
    std::vector<double> conn, offs;
    long npt = 0;
    for (auto& kv : first) {
        const geometry::Connection* c = kv.second;
        auto t = cellTypes.find(kv.first);
        if (t == cellTypes.end()) LBIB_FAIL("domainIdentifier not found in the cellTrackerMap.");
        do {
            pts.push_back(c->second->x);
            pts.push_back(c->second->y);
            pts.push_back(0.0);
            dom.push_back(c->domain);
            ct.push_back(t->second);
            std::string s = "\t";
            for (auto& b : c->descriptor)
                for (auto& cde : b.second) s += b.first + "\t" + cde + "\t";
            for (char ch : s) desc.push_back((double)(unsigned char)ch);
            desc.push_back(0.0);
            conn.push_back((double)npt++);
            c = c->second->out;
            if (!c) LBIB_FAIL("open membrane ring");
        } while (c != kv.second);
        offs.push_back((double)npt);
    }
    std::ofstream o(dirname_of(vtm) + "/" + rel);
    if (!o.is_open()) LBIB_FAIL("cannot write " + rel);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "  <PolyData>\n    <Piece NumberOfPoints=\"" << npt
      << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"" << offs.size()
      << "\">\n      <PointData>\n";
    write_array(o, "Float64", "domain_identifier", 1, dom);
    write_array(o, "Float64", "cell_type", 1, ct);
    o << "        <Array type=\"String\" Name=\"boundarydescriptors\" format=\"ascii\">";
    write_numbers(o, desc);
    o << "        </Array>\n      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n";
    write_array(o, "Float64", "Points", 3, pts);
    o << "      </Points>\n      <Polys>\n";
    write_array(o, "Int64", "connectivity", 1, conn);
    write_array(o, "Int64", "offsets", 1, offs);
    o << "      </Polys>\n    </Piece>\n  </PolyData>\n</VTKFile>\n";
    o.close();
    append_block(vtm, rel);
}

// Geometry::loadGeometryVTK (Geometry.cpp:162-306) into `geo`.
void load_geometry(const std::string& file, geometry::Geometry& geo) {
    struct stat st;
    if (stat(file.c_str(), &st) != 0) LBIB_FAIL("Cannot find the geometry.vtm file.");
    std::vector<std::string> blocks;
    if (file.size() >= 4 && file.compare(file.size() - 4, 4, ".vtp") == 0) blocks.push_back(file);
    else blocks = vtm_blocks(file);
    int found = 0;
    PolyData pd;
    for (auto& b : blocks) {
        if (b.size() < 4 || b.compare(b.size() - 4, 4, ".vtp") != 0) continue;
        PolyData cand = read_vtp(b);
        if (cand.npolys > 0) {
            pd = std::move(cand);
            found++;
        }
    }
    if (found != 1) LBIB_FAIL("more or less than one vtk polygon block found in input geometry file found.");
    const Array* da = pd.find("domain_identifier");
    if (!da) LBIB_FAIL("domain_identifier array not found in vtk input file.");
    const Array* ca = pd.find("cell_type");
    if (!ca) LBIB_FAIL("cell_type array not found in vtk input file.");
    const Array* ba = pd.find("boundarydescriptors");
    if (!ba) LBIB_FAIL("boundarydescriptor array not found in vtk input file");
    if ((long)da->num.size() != pd.npoints || (long)ca->num.size() != pd.npoints ||
        (long)ba->str.size() != pd.npoints)
        LBIB_FAIL("point attribute arrays do not match the number of points");
    if (!geo.getGeometryNodes().empty() || !geo.getConnections().empty()) LBIB_FAIL("Only load the geometry once.");
    auto& types = geo.cellTypeTrackerMap();
    types[0] = 0;   // fluid default
    for (long k = 0; k < pd.npoints; k++) geo.addGeometryNode((unsigned)k, pd.points[3 * k], pd.points[3 * k + 1]);
    std::map<std::string, std::vector<std::string>> connectionType;
    long start = 0;
    for (long i = 0; i < pd.npolys; i++) {
        const long end = pd.offs[i];
        if (end <= start || end > (long)pd.conn.size()) LBIB_FAIL("bad polygon offsets in vtk input file");
        const long* ids = pd.conn.data() + start;
        const long numIds = end - start;
        const unsigned domainId = (unsigned)da->num[ids[0]];   // the first point's value, as the reference
        types[domainId] = (unsigned)ca->num[ids[0]];
        for (long k = 0; k < numIds - 1; k++) {
            std::stringstream ss(ba->str[ids[k]]);
            connectionType.clear();
            std::string t1, t2;
            while (ss >> t1 >> t2) connectionType[t1].push_back(t2);
            geo.addConnection((unsigned)ids[k], (unsigned)ids[k + 1], domainId, connectionType);
        }
        // the closing connection takes the descriptor parsed last (Geometry.cpp:283-286)
        geo.addConnection((unsigned)ids[numIds - 1], (unsigned)ids[0], domainId, connectionType);
        start = end;
    }
}

}  // namespace vtkxml
}  // namespace LbmLib
