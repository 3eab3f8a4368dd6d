/*
 * MultiCellDS snapshots (BioFVM_MultiCellDS.cpp:394-720,
 * save_microenvironment_to_MultiCellDS_xml_pugi): an XML document describing
 * the run (metadata), the mesh (coordinates and voxels, or a mesh MAT file),
 * the densities (names, units, diffusion constants, decay rates) and the data
 * (a MAT v4 file packed on the GPU by Microenvironment::write_to_matlab, or
 * one CSV line per voxel).
 *
 * As in the reference, one document lives per process: the first call builds
 * it from the microenvironment given; later calls only update the current
 * time, the run time and the data (and write the files under the new base
 * name), leaving the mesh and density descriptions as first written.
 *
 * The document is written in the layout of the reference's XML library with
 * its default options: an `<?xml version="1.0"?>` declaration, one element
 * per line indented by tabs, text content on the element's own line,
 * attributes in insertion order with `&`, `<`, `>` and `"` escaped.  Numbers
 * use the reference's printf formats (%f; %.7e in CSV data).
 */
#include "BioFVM.h"

#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <utility>

namespace BioFVM {

MultiCellDS_Metadata default_metadata;

MultiCellDS_Metadata::MultiCellDS_Metadata()
    : MultiCellDS_type("generic"), user("Unspecified"), program("BioFVM"), program_version("1.0.4"),
      bounding_box(6, 0.0), spatial_units("microns"), time_units("min"), runtime_time_units("seconds"),
      current_time(0.0), current_runtime(0.0)
{
}

namespace {

bool g_mesh_as_matlab = false;
bool g_data_as_matlab = true;

// A minimal element tree: name, attributes in order, text, children.
struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    bool has_text = false;
    std::vector<std::unique_ptr<XNode>> kids;

    XNode* add(const std::string& n)
    {
        kids.emplace_back(new XNode);
        kids.back()->name = n;
        return kids.back().get();
    }
    XNode* attr(const std::string& k, const std::string& v)
    {
        for (auto& a : attrs)
            if (a.first == k) {
                a.second = v;
                return this;
            }
        attrs.emplace_back(k, v);
        return this;
    }
    XNode* set_text(const std::string& t)
    {
        text = t;
        has_text = true;
        return this;
    }
    XNode* child(const std::string& n)
    {
        for (auto& k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
};

std::unique_ptr<XNode> g_doc;  // the process's document (root element)

std::string escape(const std::string& s, bool attribute)
{
    std::string o;
    for (char c : s) {
        if (c == '&') o += "&amp;";
        else if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (c == '"' && attribute) o += "&quot;";
        else o += c;
    }
    return o;
}

void emit(std::ostream& os, const XNode& n, int depth)
{
    const std::string tabs(depth, '\t');
    os << tabs << '<' << n.name;
    for (const auto& a : n.attrs) os << ' ' << a.first << "=\"" << escape(a.second, true) << '"';
    if (n.kids.empty() && !n.has_text) {
        os << " />\n";
        return;
    }
    os << '>';
    if (n.kids.empty()) {
        os << escape(n.text, false) << "</" << n.name << ">\n";
        return;
    }
    os << '\n';
    if (n.has_text) os << tabs << '\t' << escape(n.text, false) << '\n';
    for (const auto& k : n.kids) emit(os, *k, depth + 1);
    os << tabs << "</" << n.name << ">\n";
}

std::string fmt(const char* f, double v)
{
    char b[64];
    std::snprintf(b, sizeof b, f, v);
    return b;
}

std::string csv(const std::vector<double>& v, const char* f)
{
    std::string o;
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) o += ',';
        o += fmt(f, v[i]);
    }
    return o;
}

std::string box(const std::vector<double>& b)
{
    char s[512];
    std::snprintf(s, sizeof s, "[%f %f %f %f %f %f]", b[0], b[1], b[2], b[3], b[4], b[5]);
    return s;
}

std::string data_file(const std::string& base)
{
    return base + "_microenvironment0.mat";  // the first (only) microenvironment
}

void fill_data(XNode* data, Microenvironment& M, const std::string& base)
{
    if (g_data_as_matlab) {
        M.write_to_matlab(data_file(base));
        data->attr("type", "matlab")->set_text(data_file(base));
        return;
    }
    data->attr("type", "xml");
    for (int j = 0; j < M.number_of_voxels(); ++j)
        data->add("density_vector")
            ->attr("voxel_ID", std::to_string(M.mesh.voxels[j].mesh_index))
            ->attr("delimiter", ",")
            ->set_text(csv(M.density_vector(j), "%.7e"));
}

// The whole document, as the reference's first call builds it.
std::unique_ptr<XNode> build(Microenvironment& M, const std::string& base)
{
    const MultiCellDS_Metadata& md = default_metadata;
    std::unique_ptr<XNode> root(new XNode);
    root->name = "MultiCellDS";
    root->attr("version", "0.3.0 beta")->attr("type", "snapshot/simulation");
    XNode* meta = root->add("metadata");
    meta->add("user")->set_text(md.user);
    meta->add("program")->attr("version", md.program_version)->set_text(md.program);
    meta->add("bounding_box")->attr("type", "axis-aligned")->attr("units", md.spatial_units)->set_text(
        box(md.bounding_box));
    meta->add("current_time")->attr("units", md.time_units)->set_text(fmt("%f", md.current_time));
    meta->add("current_runtime")->attr("units", md.runtime_time_units)->set_text(fmt("%f", md.current_runtime));

    XNode* scale = root->add("microenvironment")->add("scale");
    scale->attr("name", M.name);
    Cartesian_Mesh& mesh = M.mesh;
    XNode* xm = scale->add("mesh");
    xm->attr("type", mesh.Cartesian_mesh ? "Cartesian" : "general")
        ->attr("uniform", mesh.uniform_mesh ? "true" : "false")
        ->attr("regular", mesh.regular_mesh ? "true" : "false")
        ->attr("units", mesh.units);
    xm->add("bounding_box")->attr("type", "axis-aligned")->attr("units", mesh.units)->set_text(box(mesh.bounding_box));
    if (mesh.Cartesian_mesh) {
        xm->add("x_coordinates")->set_text(csv(mesh.x_coordinates, "%f"))->attr("delimiter", ",");
        xm->add("y_coordinates")->set_text(csv(mesh.y_coordinates, "%f"))->attr("delimiter", ",");
        xm->add("z_coordinates")->set_text(csv(mesh.z_coordinates, "%f"))->attr("delimiter", ",");
    }
    XNode* vox = xm->add("voxels");
    if (!g_mesh_as_matlab) {
        vox->attr("type", "xml");
        for (const Voxel& v : mesh.voxels) {
            XNode* e = vox->add("voxel")->attr("ID", std::to_string(v.mesh_index))->attr("type", "cube");
            e->add("center")->attr("delimiter", ",")->set_text(csv(v.center, "%f"));
            e->add("volume")->set_text(fmt("%f", v.volume));
        }
    } else {
        const std::string f = base + "_mesh0.mat";
        mesh.write_to_matlab(f);
        vox->attr("type", "matlab")->set_text(f);
    }
    XNode* dens = scale->add("densities");
    for (int j = 0; j < M.number_of_densities(); ++j) {
        XNode* d = dens->add("density")->attr("name", M.density_names[j])->attr("units", M.density_units[j]);
        d->add("diffusion_constant")->set_text(fmt("%f", M.diffusion_coefficients[j]));
        d->add("decay_rate")->set_text(fmt("%f", M.decay_rates[j]));
    }
    fill_data(scale->add("data"), M, base);
    return root;
}

// A later call: current time and run time, then the data (BioFVM_MultiCellDS.cpp:401-468).
void update(XNode* root, Microenvironment& M, const std::string& base)
{
    XNode* meta = root->child("metadata");
    if (XNode* n = meta ? meta->child("current_time") : nullptr) n->set_text(fmt("%f", default_metadata.current_time));
    if (XNode* n = meta ? meta->child("current_runtime") : nullptr)
        n->set_text(fmt("%f", default_metadata.current_runtime));
    XNode* me = root->child("microenvironment");
    XNode* scale = me ? me->child("scale") : nullptr;
    XNode* data = scale ? scale->child("data") : nullptr;
    if (!data) return;
    if (g_data_as_matlab) {
        M.write_to_matlab(data_file(base));
        data->attr("type", "matlab");
        if (data->kids.empty()) data->set_text(data_file(base));  // a CSV-data document keeps its elements
        return;
    }
    data->attr("type", "xml");
    // the existing per-voxel lines, in order, as far as they go
    int j = 0;
    for (auto& k : data->kids)
        if (k->name == "density_vector" && j < M.number_of_voxels()) k->set_text(csv(M.density_vector(j++), "%.7e"));
}

}  // namespace

void set_save_biovm_mesh_as_matlab(bool newvalue) { g_mesh_as_matlab = newvalue; }
void set_save_biovm_data_as_matlab(bool newvalue) { g_data_as_matlab = newvalue; }

void save_microenvironment_to_MultiCellDS_xml_pugi(std::string filename_base, Microenvironment& M)
{
    std::cout << "Writing data to file " << filename_base << ".xml ... ";
    if (g_doc) update(g_doc.get(), M, filename_base);
    else g_doc = build(M, filename_base);
    std::ofstream os(filename_base + ".xml", std::ios::binary);
    os << "<?xml version=\"1.0\"?>\n";
    emit(os, *g_doc, 0);
    std::cout << "done!" << std::endl;
}

// BioFVM_mesh.cpp:350-367 (General_Mesh::write_to_matlab): MAT v4, a 4 x voxels
// FP64 matrix named "mesh", one column (x, y, z, volume) per voxel.
void Cartesian_Mesh::write_to_matlab(std::string filename)
{
    FILE* fp = std::fopen(filename.c_str(), "wb");
    if (!fp) {
        std::cout << "Error: could not create file " << filename << "!" << std::endl;
        return;
    }
    const char name[] = "mesh";
    const unsigned header[5] = {0u, 4u, (unsigned)voxels.size(), 0u, (unsigned)(sizeof(name) - 1)};
    std::fwrite(header, sizeof(unsigned), 5, fp);
    std::fwrite(name, 1, sizeof(name) - 1, fp);
    for (const Voxel& v : voxels) {
        const double col[4] = {v.center[0], v.center[1], v.center[2], v.volume};
        std::fwrite(col, sizeof(double), 4, fp);
    }
    std::fclose(fp);
}

}  // namespace BioFVM
