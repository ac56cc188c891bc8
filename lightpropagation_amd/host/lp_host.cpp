// lp_host.cpp -- the host layer of lp_host.hpp and its C ABI
// (include/lightprop_host.h).
#include "lp_host.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>

#include "lightprop_host.h"

namespace lph {
namespace {

// ---------------------------------------------------------------------------
// minimal XML reader: elements, character data, attributes skipped;
// declarations, comments and processing instructions ignored.  Enough for
// the configuration files ReadXML parses (src/IO/ReadXML.cpp).
// ---------------------------------------------------------------------------
struct XNode {
    std::string name, text;
    std::vector<XNode> kids;
    const XNode* child(const char* n) const
    {
        for (const XNode& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
};

bool parse_xml(const std::string& s, XNode& root, std::string& err)
{
    std::vector<XNode*> stack;
    root = XNode();
    root.name = "#document";
    stack.push_back(&root);
    size_t i = 0, n = s.size();
    while (i < n) {
        size_t lt = s.find('<', i);
        if (lt == std::string::npos) lt = n;
        if (lt > i) stack.back()->text.append(s, i, lt - i);
        if (lt >= n) break;
        if (s.compare(lt, 4, "<!--") == 0) {
            size_t e = s.find("-->", lt + 4);
            if (e == std::string::npos) return err = "unterminated comment", false;
            i = e + 3;
            continue;
        }
        if (s.compare(lt, 2, "<?") == 0 || s.compare(lt, 2, "<!") == 0) {
            size_t e = s.find('>', lt);
            if (e == std::string::npos) return err = "unterminated declaration", false;
            i = e + 1;
            continue;
        }
        size_t gt = s.find('>', lt);
        if (gt == std::string::npos) return err = "unterminated tag", false;
        std::string tag = s.substr(lt + 1, gt - lt - 1);
        i = gt + 1;
        if (!tag.empty() && tag[0] == '/') {
            std::string name = tag.substr(1);
            while (!name.empty() && isspace((unsigned char)name.back())) name.pop_back();
            if (stack.size() < 2 || stack.back()->name != name) return err = "mismatched </" + name + ">", false;
            stack.pop_back();
            continue;
        }
        bool self = !tag.empty() && tag.back() == '/';
        if (self) tag.pop_back();
        size_t e = 0;
        while (e < tag.size() && !isspace((unsigned char)tag[e])) ++e;
        XNode node;
        node.name = tag.substr(0, e);
        stack.back()->kids.push_back(node);
        if (!self) stack.push_back(&stack.back()->kids.back());
    }
    if (stack.size() != 1) return err = "unclosed element <" + stack.back()->name + ">", false;
    return true;
}

// atof / atoi of the element's character data, as ReadXML does
double num(const XNode& x) { return atof(x.text.c_str()); }
std::string trimmed(const std::string& t)
{
    size_t a = 0, b = t.size();
    while (a < b && isspace((unsigned char)t[a])) ++a;
    while (b > a && isspace((unsigned char)t[b - 1])) --b;
    return t.substr(a, b - a);
}

bool is_word(char c) { return isalnum((unsigned char)c) || c == '_'; }

// Whole-line matches of the reference's regular expressions
// (ImportExport.cpp:967-970, 398): "^o (\w+)", "^f (.+)", "^g (\d+)$",
// "^g (\d+)\_(\w+)".
bool match_o(const std::string& l, std::string& name)
{
    if (l.size() < 3 || l[0] != 'o' || l[1] != ' ') return false;
    for (size_t k = 2; k < l.size(); ++k)
        if (!is_word(l[k])) return false;
    name = l.substr(2);
    return true;
}
bool match_f(const std::string& l) { return l.size() >= 3 && l[0] == 'f' && l[1] == ' '; }
bool match_g(const std::string& l, int& id, std::string& word, bool& has_word)
{
    if (l.size() < 3 || l[0] != 'g' || l[1] != ' ') return false;
    size_t k = 2;
    while (k < l.size() && isdigit((unsigned char)l[k])) ++k;
    if (k == 2) return false;
    id = atoi(l.substr(2, k - 2).c_str());
    if (k == l.size()) {
        has_word = false;
        return true;
    }
    if (l[k] != '_' || k + 1 >= l.size()) return false;
    for (size_t m = k + 1; m < l.size(); ++m)
        if (!is_word(l[m])) return false;
    word = l.substr(k + 1);
    has_word = true;
    return true;
}
// "^(\d+\.?\d*) (\d+\.?\d*)$" of the emitter file (ImportExport.cpp:398)
bool match_number(const std::string& s)
{
    size_t k = 0;
    while (k < s.size() && isdigit((unsigned char)s[k])) ++k;
    if (k == 0) return false;
    if (k < s.size() && s[k] == '.') ++k;
    while (k < s.size() && isdigit((unsigned char)s[k])) ++k;
    return k == s.size();
}

bool file_exists(const std::string& p)
{
    struct stat st;
    return !p.empty() && stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}
std::string basename_of(const std::string& p)
{
    size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}
std::string dirname_of(const std::string& p)
{
    size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

// %g, the default formatting of the reference's ofstream << float
std::string g(double v)
{
    char b[64];
    snprintf(b, sizeof b, "%g", v);
    return b;
}

// float helpers with the reference's operation order (Emitter.cpp:475-541)
void normalize3(float a[3])
{
    float f = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] = a[0] / f;
    a[1] = a[1] / f;
    a[2] = a[2] / f;
}
void cross3(const float a[3], const float b[3], float r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

// ---------------------------------------------------------------------------
// Image
// ---------------------------------------------------------------------------
void Image::GetExtent(int e[6]) const { std::memcpy(e, ext, sizeof ext); }
float Image::GetScalarComponentAsFloat(int x, int y, int, int c) const
{
    return data[((size_t)(y - ext[2]) * nx() + (x - ext[0])) * ncomp + c];
}
void Image::SetScalarComponentFromFloat(int x, int y, int, int c, float v)
{
    data[((size_t)(y - ext[2]) * nx() + (x - ext[0])) * ncomp + c] = v;
}

// ---------------------------------------------------------------------------
// ImportExport
// ---------------------------------------------------------------------------
int ImportExport::Fail(const std::string& msg)
{
    error = msg;
    return 1;
}

std::string ImportExport::Resolve(const std::string& path) const
{
    if (file_exists(path)) return path;
    std::string alt = dirname_of(configfile) + "/" + basename_of(path);
    return file_exists(alt) ? alt : path;
}

// ReadXML::ParseConfig and its element handlers (ReadXML.cpp:93-168,
// 598-622, 663-871, 950-1010), then ImportExport::ImportModel (:257-262).
int ImportExport::ParseFile()
{
    std::ifstream f(configfile);
    if (!f) return Fail("cannot open configuration " + configfile);
    std::stringstream ss;
    ss << f.rdbuf();
    XNode doc;
    std::string err;
    if (!parse_xml(ss.str(), doc, err)) return Fail(configfile + ": " + err);
    if (doc.kids.empty()) return Fail(configfile + ": no root element");
    const XNode& root = doc.kids.front();
    for (const XNode& e : root.kids) {
        if (e.name == "modelfile") {
            modelfile = trimmed(e.text);
        } else if (e.name == "saveLogFiles") {
            saveFile = atoi(e.text.c_str()) != 0;
        } else if (e.name == "saveDir") {
            saveDir = trimmed(e.text);
        } else if (e.name == "model") {
            for (const XNode& m : e.kids) {
                if (m.name == "modelfile") {
                    modelfile = trimmed(m.text);
                } else if (m.name == "PixelPerUnit") {
                    ppu = (float)num(m);
                } else if (m.name == "dimensions") {
                    for (const XNode& d : m.kids) {
                        if (d.name == "x") dims[0] = (float)num(d);
                        else if (d.name == "y") dims[1] = (float)num(d);
                        else if (d.name == "z") dims[2] = (float)num(d);
                    }
                } else if (m.name == "tissueList") {
                    // arrays sized by the list's element count, indexed by ID
                    size_t nt = m.kids.size();
                    refraction.assign(nt, 0.0f);
                    absorption.assign(nt, 0.0f);
                    anisotropy.assign(nt, 0.0f);
                    scatter.assign(nt, 0.0f);
                    tissueNames.assign(nt, "");
                    for (const XNode& t : m.kids) {
                        if (t.name != "tissue") continue;
                        int id = 0;
                        for (const XNode& q : t.kids) {
                            if (q.name == "ID") id = atoi(q.text.c_str());
                            if (id < 0 || (size_t)id >= nt) return Fail("tissue ID out of range in " + configfile);
                            if (q.name == "name") tissueNames[id] = trimmed(q.text);
                            else if (q.name == "anisotropy") anisotropy[id] = (float)num(q);
                            else if (q.name == "mua") absorption[id] = (float)num(q);
                            else if (q.name == "mus") scatter[id] = (float)num(q);
                            else if (q.name == "refraction") refraction[id] = (float)num(q);
                        }
                    }
                }
            }
        } else if (e.name == "Emittermodel") {
            for (const XNode& m : e.kids) {
                if (m.name == "file") {
                    emitterFile = trimmed(m.text);
                } else if (m.name == "distance") {
                    emitterDistance = (float)num(m);
                } else if (m.name == "position" || m.name == "heading") {
                    std::vector<float>& v = m.name == "position" ? emitterPoint : emitterHeading;
                    for (const XNode& d : m.kids) {
                        if (d.name == "x") v[0] = (float)num(d);
                        else if (d.name == "y") v[1] = (float)num(d);
                        else if (d.name == "z") v[2] = (float)num(d);
                    }
                }
            }
        } else if (e.name == "SimulationParams") {
            for (const XNode& m : e.kids) {
                if (m.name == "numberOfPhotons") numPhotons = atoi(m.text.c_str());
                else if (m.name == "numPhotonsInteractions") numPhotonInteractions = atoi(m.text.c_str());
                else if (m.name == "chance") chance = (float)num(m);
                else if (m.name == "weightThreshold") wThreshold = (float)num(m);
            }
        } else if (e.name == "Logs") {
            for (const XNode& m : e.kids) {
                if (m.name == "saveLogFiles") {
                    saveFile = atoi(m.text.c_str()) != 0;
                } else if (m.name == "fetalID") {
                    fetalID = atoi(m.text.c_str());
                } else if (m.name == "DetectionGrid") {
                    detGrid.assign(6, 1.0f);  // ExtractLogsDetectionGrid (ReadXML.cpp:848-870)
                    for (const XNode& d : m.kids) {
                        if (d.name == "x") detGrid[0] = (float)num(d);
                        else if (d.name == "y") detGrid[1] = (float)num(d);
                        else if (d.name == "z") detGrid[2] = (float)num(d);
                        else if (d.name == "r") detGrid[3] = (float)num(d);
                        else if (d.name == "phi") detGrid[4] = (float)num(d);
                        else if (d.name == "rho") detGrid[5] = (float)num(d);
                    }
                } else if (m.name == "filenames") {
                    for (const XNode& d : m.kids) {
                        std::string v = trimmed(d.text);
                        if (d.name == "surface") surfaceFilename = v;
                        else if (d.name == "fetal") fetalFilename = v;
                        else if (d.name == "transmission") transmissionFilename = v;
                        else if (d.name == "absorbtionXZ") recordAxzFilename = v;
                        else if (d.name == "absorbtionYZ") recordAyzFilename = v;
                    }
                }
            }
        }
    }
    if (modelfile.empty()) return Fail(configfile + ": no <modelfile>");
    int rc = ImportObjModel(Resolve(modelfile));
    if (rc) return rc;
    if (!emitterFile.empty()) return ImportEmitterModel(Resolve(emitterFile));
    angles = {0.0f};
    rri = {1.0f};
    return 0;
}

// vtkOBJReader's vertices and faces in file order, the first three points of
// each face, and GenMediumIDsFromObjModel's (inner, outer) ids
// (ImportExport.cpp:952-1053).  Cell normals are the unit cross product of
// the face edges (P1 - P0) x (P2 - P0), in double.
int ImportExport::ImportObjModel(const std::string& fname)
{
    std::ifstream f(fname);
    if (!f) return Fail("cannot open model " + fname);
    std::vector<float> verts;
    std::vector<int64_t> faces;
    mesh = Mesh();
    surfaceIDs.clear();
    surfaceIDnames.clear();
    int inner = 0, outer = 0;
    bool isSurface = false;
    std::string surfName, line;
    while (std::getline(f, line)) {
        std::string name, word;
        int gid;
        bool has_word;
        if (line.compare(0, 2, "v ") == 0) {
            std::istringstream is(line.substr(2));
            std::string x, y, z;
            is >> x >> y >> z;
            verts.push_back((float)strtod(x.c_str(), nullptr));
            verts.push_back((float)strtod(y.c_str(), nullptr));
            verts.push_back((float)strtod(z.c_str(), nullptr));
        } else if (match_o(line, name)) {
            for (size_t t = 0; t < tissueNames.size(); ++t)
                if (tissueNames[t] == name) {
                    inner = (int)t;
                    break;
                }
        } else if (match_f(line)) {
            std::istringstream is(line.substr(2));
            std::string tok;
            int64_t idx[3];
            int k = 0;
            while (k < 3 && (is >> tok)) {
                long v = strtol(tok.c_str(), nullptr, 10);
                idx[k++] = v < 0 ? (int64_t)verts.size() / 3 + v : v - 1;
            }
            if (k < 3) return Fail(fname + ": face with fewer than 3 vertices");
            faces.insert(faces.end(), idx, idx + 3);
            mesh.med.push_back(inner);
            mesh.med.push_back(outer);
            if (isSurface) {
                surfaceIDs.push_back((int)(mesh.med.size() / 2 - 1));
                surfaceIDnames.push_back(surfName);
                surfName = "";
            }
        } else if (match_g(line, gid, word, has_word)) {
            outer = gid;
            isSurface = has_word;
            if (has_word) surfName = word;
        }
    }
    const int64_t nv = (int64_t)verts.size() / 3, nt = (int64_t)faces.size() / 3;
    mesh.tri.resize(9 * nt);
    mesh.nrm.resize(3 * nt);
    for (int64_t t = 0; t < nt; ++t) {
        double P[3][3];
        for (int k = 0; k < 3; ++k) {
            int64_t v = faces[3 * t + k];
            if (v < 0 || v >= nv) return Fail(fname + ": face vertex index out of range");
            for (int a = 0; a < 3; ++a) {
                mesh.tri[9 * t + 3 * k + a] = verts[3 * v + a];
                P[k][a] = verts[3 * v + a];
            }
        }
        double u[3], w[3], c[3];
        for (int a = 0; a < 3; ++a) {
            u[a] = P[1][a] - P[0][a];
            w[a] = P[2][a] - P[0][a];
        }
        c[0] = u[1] * w[2] - u[2] * w[1];
        c[1] = u[2] * w[0] - u[0] * w[2];
        c[2] = u[0] * w[1] - u[1] * w[0];
        double len = std::sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
        for (int a = 0; a < 3; ++a) mesh.nrm[3 * t + a] = (float)(c[a] / len);
    }
    return 0;
}

// ImportEmitterModel (ImportExport.cpp:393-428): "angle intensity" lines.
int ImportExport::ImportEmitterModel(const std::string& fname)
{
    std::ifstream f(fname);
    if (!f) return Fail("cannot open emitter file " + fname);
    angles.clear();
    rri.clear();
    std::string line;
    while (std::getline(f, line)) {
        size_t sp = line.find(' ');
        if (sp == std::string::npos) continue;
        std::string a = line.substr(0, sp), b = line.substr(sp + 1);
        if (!match_number(a) || !match_number(b)) continue;
        angles.push_back((float)atof(a.c_str()));
        rri.push_back((float)atof(b.c_str()));
    }
    if (angles.empty()) return Fail(fname + ": no emitter lines");
    return 0;
}

// ---------------------------------------------------------------------------
// Emitter (Emitter.cpp:186-339)
// ---------------------------------------------------------------------------
void Emitter::Calculate()
{
    // CalculatePointOfOrigin
    position.assign(3, 0.0f);
    for (int k = 0; k < 3; ++k) position[k] = -heading[k] * distance + p0[k];
    // CalculateStartVectorArray: the heading, rotated by each non-zero angle
    // (RotateVector, :389-473, degrees).  The reference's three-operand
    // VecAdd adds c[0] to all three components (:519-529); the rotated
    // vector here is the sum it means.
    startVector.clear();
    for (float phi : angles) {
        if (phi == 0.0f) {
            startVector.insert(startVector.end(), heading.begin(), heading.end());
            continue;
        }
        const float* v = heading.data();
        float ph = (float)(phi * (M_PI / 180.0));
        float cp = std::cos(ph);
        float sp = std::sqrt(1 - cp * cp);
        float e[3] = {0, 0, 0};
        if (v[2] != 0) {
            e[0] = 1;
            e[1] = 1;
            e[2] = -(v[0] + v[1]) / v[2];
        } else if (v[1] != 0) {
            e[0] = 1;
            e[1] = -(v[0] + v[2]) / v[1];
            e[2] = 1;
        } else if (v[0] != 0) {
            e[0] = -(v[1] + v[2]) / v[0];
            e[1] = 1;
            e[2] = 1;
        }
        normalize3(e);
        float nx[3], nxe[3];
        cross3(e, v, nx);
        cross3(nx, e, nxe);
        float d = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
        float r[3];
        for (int k = 0; k < 3; ++k) r[k] = d * e[k] + cp * nxe[k] + sp * nx[k];
        normalize3(r);
        startVector.insert(startVector.end(), r, r + 3);
    }
    // CalculateStartVectorOccurrences
    occ.clear();
    float sum = 0;
    for (float x : rri) sum += x;
    int64_t s1 = 0;
    for (float x : rri) {
        int o = (int)((float)numberOfPhotons * x / sum);
        s1 += o;
        occ.push_back(o);
    }
    if (!occ.empty() && s1 < numberOfPhotons) occ[0] += (int)(numberOfPhotons - s1);
    // basis of the emitter frame
    float z[3] = {heading[0], heading[1], heading[2]}, a[3] = {0, 0, 0}, y[3], x[3];
    if (z[2] != 0) {
        a[0] = 1;
        a[1] = 1;
        a[2] = -(z[0] + z[1]) / z[2];
    } else if (z[1] != 0) {
        a[0] = 1;
        a[1] = -z[0] / z[1];
    } else if (z[0] != 0) {
        a[1] = 1;
    }
    normalize3(a);
    normalize3(z);
    cross3(z, a, y);
    normalize3(y);
    cross3(y, z, x);
    normalize3(x);
    ex.assign(x, x + 3);
    ey.assign(y, y + 3);
    ez.assign(z, z + 3);
}

void Emitter::GetEmittermodel(std::vector<float>& pStart, std::vector<float>& vStart, std::vector<int>& vStartOcc)
{
    Calculate();
    pStart = position;
    vStart = startVector;
    vStartOcc = occ;
}

void Emitter::GetBasisVectors(std::vector<float>& eX, std::vector<float>& eY, std::vector<float>& eZ)
{
    if (ex.empty()) Calculate();
    eX = ex;
    eY = ey;
    eZ = ez;
}

// ---------------------------------------------------------------------------
// CalculatePropagation
// ---------------------------------------------------------------------------
CalculatePropagation::~CalculatePropagation()
{
    if (ctx) lp_destroy(ctx);
}

int CalculatePropagation::Fail(const std::string& msg)
{
    error = msg;
    return LP_FAILURE;
}

void CalculatePropagation::SetTissueProperties(const std::vector<float>& refractionInd,
                                               const std::vector<float>& absorbtionCoeffs,
                                               const std::vector<float>& anisotropies,
                                               const std::vector<float>& scatterCoeffs)
{
    // SetTissueProperties (CalculatePropagation.cpp:427-451): (n, mua, mus, g)
    size_t n = refractionInd.size();
    nmuasg.assign(4 * n, 0.0f);
    for (size_t i = 0; i < n; ++i) {
        nmuasg[4 * i] = refractionInd[i];
        nmuasg[4 * i + 1] = i < absorbtionCoeffs.size() ? absorbtionCoeffs[i] : 0.0f;
        nmuasg[4 * i + 2] = i < scatterCoeffs.size() ? scatterCoeffs[i] : 0.0f;
        nmuasg[4 * i + 3] = i < anisotropies.size() ? anisotropies[i] : 0.0f;
    }
}

int CalculatePropagation::Init()
{
    if (ctx) {
        lp_destroy(ctx);
        ctx = nullptr;
    }
    int rc = lp_create(&ctx, device);
    if (rc) return Fail(ctx ? lp_last_error(ctx) : "lp_create failed");
    auto chk = [&](int r) {
        if (r && error.empty()) error = lp_last_error(ctx);
        return r;
    };
    if (dims.size() < 3 || detGrid.size() < 6) return Fail("dimensions and detection grid must be set before Init");
    float pb[12];
    for (int k = 0; k < 3; ++k) {
        pb[k] = poi.at(k);
        pb[3 + k] = eX.at(k);
        pb[6 + k] = eY.at(k);
        pb[9 + k] = eZ.at(k);
    }
    error.clear();
    if (chk(lp_set_mesh(ctx, mesh.tri.data(), mesh.nrm.data(), mesh.med.data(), mesh.size()))) return LP_FAILURE;
    if (chk(lp_set_materials(ctx, nmuasg.data(), (int32_t)(nmuasg.size() / 4), fetalID))) return LP_FAILURE;
    if (chk(lp_set_source(ctx, pStart.data(), vStart.data(), vStartOcc.data(), (int32_t)vStartOcc.size(), pb)))
        return LP_FAILURE;
    if (chk(lp_set_grid(ctx, dims.data(), detGrid.data()))) return LP_FAILURE;
    if (chk(lp_set_params(ctx, chance, wThreshold, factor, seed, rng, 0))) return LP_FAILURE;
    const uint32_t mask = (1u << LP_TALLY_SFT) | (1u << LP_TALLY_RTD) | (1u << LP_TALLY_ARZ) |
                          (1u << LP_TALLY_PHIRZ) | (1u << LP_TALLY_REGION) | (1u << LP_TALLY_MISC);
    if (chk(lp_set_tallies(ctx, mask, 0))) return LP_FAILURE;
    if (chk(lp_reset_tallies(ctx))) return LP_FAILURE;
    return LP_OK;
}

int CalculatePropagation::Propagation()
{
    if (!ctx) return Fail("Propagation before Init");
    if (lp_run(ctx, 0, numberOfPhotons)) return Fail(lp_last_error(ctx));
    lp_stats st;
    if (lp_get_stats(ctx, &st) == LP_OK) kernelMs = st.kernel_ms;
    return LP_OK;
}

void CalculatePropagation::PrintTimeSpent() const
{
    std::cout << "Time spent in propagation: " << kernelMs << " ms (" << numberOfPhotons << " photons)" << std::endl;
}

namespace {

// a tally of ctx as the reference's float image (see lph_image)
int tally_image(lp_context* ctx, int which, int component, Image& img)
{
    if (!ctx) return LP_ESTATE;
    lp_layout L;
    int rc = lp_get_layout(ctx, &L);
    if (rc) return rc;
    std::vector<int64_t> t((size_t)L.n[which]);
    if ((rc = lp_get_tally(ctx, which, t.data(), L.n[which]))) return rc;
    const double unit = which == LP_TALLY_PHIRZ ? LP_FIXED_SCALE_PHI : LP_FIXED_SCALE;
    auto val = [&](int64_t v) { return (float)((double)v / unit); };
    img = Image();
    if (which == LP_TALLY_SFT || which == LP_TALLY_RTD) {
        int nc = 3;
        if (which == LP_TALLY_SFT) {
            int ex[6] = {-(L.dimx / 2), L.dimx / 2, -(L.dimy / 2), L.dimy / 2, 0, 0};
            std::memcpy(img.ext, ex, sizeof ex);
        } else {
            int ex[6] = {0, L.dimR - 1, 0, L.dimAlpha - 1, 0, 0};
            std::memcpy(img.ext, ex, sizeof ex);
        }
        img.ncomp = component < 0 ? nc : 1;
        img.data.assign((size_t)img.nx() * img.ny() * img.ncomp, 0.0f);
        for (int y = 0; y < img.ny(); ++y)
            for (int x = 0; x < img.nx(); ++x)
                for (int c = 0; c < img.ncomp; ++c) {
                    int src_c = component < 0 ? c : component;
                    // SFT tally: (y, x, c); RTd tally: (alpha, r, c) -- the
                    // image's y is the tally's slow index in both
                    int64_t i = ((int64_t)y * img.nx() + x) * nc + src_c;
                    img.data[((size_t)y * img.nx() + x) * img.ncomp + c] = val(t[i]);
                }
        return LP_OK;
    }
    if (which == LP_TALLY_ARZ || which == LP_TALLY_PHIRZ) {
        int ex[6] = {-L.dimR, L.dimR, -L.dimz, 0, 0, 0};
        std::memcpy(img.ext, ex, sizeof ex);
        img.ncomp = 1;
        img.data.assign((size_t)img.nx() * img.ny(), 0.0f);
        // tally (cz, r + dimR), cz in [0, dimz] -> image (r, -cz)
        // (CalculatePropagation.cpp:1646-1687)
        for (int cz = 0; cz <= L.dimz; ++cz)
            for (int r = -L.dimR; r <= L.dimR; ++r)
                img.SetScalarComponentFromFloat(r, -cz, 0, 0, val(t[(int64_t)cz * L.arz_nr + (r + L.dimR)]));
        return LP_OK;
    }
    return LP_EINVAL;
}

}  // namespace

Image CalculatePropagation::GetSFTImage() const
{
    Image i;
    tally_image(ctx, LP_TALLY_SFT, -1, i);
    return i;
}
Image CalculatePropagation::GetSFTImage(int c) const
{
    Image i;
    tally_image(ctx, LP_TALLY_SFT, c, i);
    return i;
}
Image CalculatePropagation::GetRTdImage() const
{
    Image i;
    tally_image(ctx, LP_TALLY_RTD, -1, i);
    return i;
}
Image CalculatePropagation::GetRTdImage(int c) const
{
    Image i;
    tally_image(ctx, LP_TALLY_RTD, c, i);
    return i;
}
Image CalculatePropagation::GetArzImage() const
{
    Image i;
    tally_image(ctx, LP_TALLY_ARZ, 0, i);
    return i;
}
Image CalculatePropagation::GetPhirzImage() const
{
    Image i;
    tally_image(ctx, LP_TALLY_PHIRZ, 0, i);
    return i;
}

// ---------------------------------------------------------------------------
// Analysis
// ---------------------------------------------------------------------------
void Analysis::SetRTdxyLogs(const Image& img)
{
    // x outer, y inner (Analysis.cpp:447-463)
    for (int x = img.ext[0]; x <= img.ext[1]; ++x)
        for (int y = img.ext[2]; y <= img.ext[3]; ++y) {
            surfRxyLog.push_back(img.GetScalarComponentAsFloat(x, y, 0, 0));
            fetalRxyLog.push_back(img.GetScalarComponentAsFloat(x, y, 0, 1));
            transTxyLog.push_back(img.GetScalarComponentAsFloat(x, y, 0, 2));
        }
}

void Analysis::CalculateSurfaceSums()
{
    for (float v : surfRxyLog) surfRxySum += v;
    for (float v : transTxyLog) transTxySum += v;
    for (float v : fetalRxyLog) fetalRxySum += v;
}

void Analysis::PrintAnalysis() const
{
    std::cout << "################ Analysis ##################" << std::endl;
    std::cout << "Number of Photons: " << numberPhotons << std::endl;
    std::cout << "SurfRxy: " << surfRxySum << " | in %: " << surfRxySum / numberPhotons * 100 << std::endl;
    std::cout << "TransRxy: " << transTxySum << " | in %: " << transTxySum / numberPhotons * 100 << std::endl;
    std::cout << "FetalRxy: " << fetalRxySum << " | in %: " << fetalRxySum / numberPhotons * 100 << std::endl;
    std::cout << "Surf+TransRxy: " << surfRxySum + transTxySum << std::endl;
    std::cout << "############################################" << std::endl;
}

// ---------------------------------------------------------------------------
// Logs (Logs.cpp:337-402, 507-710, 1075-1110), float where the reference
// keeps float and double where it computes in double
// ---------------------------------------------------------------------------
float Logs::ScalePerRadius1D(int ir) const
{
    double dr = gridDims[3];
    float scale1 = (float)(2.0 * M_PI * dr * dr * numberPhotons);
    return (float)(1.0 / ((ir + 0.5) * scale1));
}

float Logs::ScalePerAngle1D(int ia) const
{
    double da = gridDims[5] * M_PI / 180;
    float scale1 = (float)(2.0 * M_PI * da * numberPhotons);
    return (float)(1.0 / (std::sin((ia + 0.5) * da) * scale1));
}

float Logs::Scale2D(int ir, int ia) const
{
    double dr = gridDims[3];
    double da = gridDims[5] * M_PI / 180;
    float scale1 = (float)(4.0 * M_PI * M_PI * dr * std::sin(da / 2) * dr * numberPhotons);
    return (float)(1.0 / ((ir + 0.5) * std::sin(2.0 * (ia + 0.5) * da) * scale1));
}

void Logs::ReflectancePerRadius(const Image& im, std::vector<float>& c0, std::vector<float>& c1) const
{
    c0.clear();
    c1.clear();
    for (int r = im.ext[0]; r <= im.ext[1]; ++r) {
        float refl = 0;
        for (int a = im.ext[2]; a <= im.ext[3]; ++a) refl += im.GetScalarComponentAsFloat(r, a, 0, 0);
        c0.push_back(r * gridDims[3]);
        c1.push_back(refl * ScalePerRadius1D(std::abs(r)));
    }
}

void Logs::ReflectancePerAngle(const Image& im, std::vector<float>& c0, std::vector<float>& c1) const
{
    c0.clear();
    c1.clear();
    for (int n = im.ext[2]; n <= im.ext[3]; ++n) {
        float refl = 0;
        for (int m = im.ext[0]; m <= im.ext[1]; ++m) refl += im.GetScalarComponentAsFloat(m, n, 0, 0);
        c0.push_back(n * gridDims[5]);
        c1.push_back(refl * ScalePerAngle1D(n));
    }
}

void Logs::FluencePerZ(const Image& im, std::vector<float>& c0, std::vector<float>& c1) const
{
    c0.clear();
    c1.clear();
    for (int z = im.ext[2]; z <= im.ext[3]; ++z) {
        float fz = 0;
        for (int r = im.ext[0]; r <= im.ext[1]; ++r) fz += im.GetScalarComponentAsFloat(r, z, 0, 0);
        c0.push_back(std::abs(z) * gridDims[2]);
        c1.push_back(fz / (gridDims[2] * (float)numberPhotons));
    }
}

namespace {
std::string stem(const std::string& fname) { return fname.substr(0, fname.find_last_of('.')); }
int write_columns(const std::string& path, const std::vector<float>& c0, const std::vector<float>& c1)
{
    std::ofstream f(path);
    if (!f) return LP_FAILURE;
    for (size_t i = 0; i < c0.size(); ++i) f << g(c0[i]) << "\t" << g(c1[i]) << "\n";
    return f ? LP_OK : LP_FAILURE;
}
}  // namespace

int Logs::SaveSurfaceArr(const std::string& fname, const Image& im) const
{
    std::ofstream f(saveDir + fname);
    if (!f) return LP_FAILURE;
    for (int n = im.ext[2]; n <= im.ext[3]; ++n) {
        for (int m = im.ext[0]; m <= im.ext[1]; ++m) f << g(im.GetScalarComponentAsFloat(m, n, 0, 0)) << "\t";
        f << "\n";
    }
    return f ? LP_OK : LP_FAILURE;
}

int Logs::SaveRecordRd(const std::string& fname, const Image& im) const
{
    std::ofstream f(saveDir + stem(fname) + "_Rd.txt");
    if (!f) return LP_FAILURE;
    for (int n = im.ext[2]; n <= im.ext[3]; ++n) {
        for (int m = im.ext[0]; m <= im.ext[1]; ++m)
            f << g(im.GetScalarComponentAsFloat(m, n, 0, 0) * Scale2D(m, n)) << "\t";
        f << "\n";
    }
    return f ? LP_OK : LP_FAILURE;
}

int Logs::CalcReflectancePerRadius(const std::string& fname, const Image& im) const
{
    std::vector<float> a, b;
    ReflectancePerRadius(im, a, b);
    return write_columns(saveDir + stem(fname) + "_perRadius.txt", a, b);
}

int Logs::CalcReflectancePerAngle(const std::string& fname, const Image& im) const
{
    std::vector<float> a, b;
    ReflectancePerAngle(im, a, b);
    return write_columns(saveDir + stem(fname) + "_perAngle.txt", a, b);
}

int Logs::FluenceZ(const std::string& fname, const Image& im) const
{
    std::vector<float> a, b;
    FluencePerZ(im, a, b);
    return write_columns(saveDir + stem(fname) + "_perZ.txt", a, b);
}

}  // namespace lph

// ===========================================================================
// C ABI (include/lightprop_host.h)
// ===========================================================================
using namespace lph;

struct lph_scene {
    ImportExport io;
    std::vector<float> pStart, vStart, eX, eY, eZ;
    std::vector<int> vStartOcc;
};

namespace {
thread_local std::string g_err;
int herr(int rc, const std::string& msg)
{
    g_err = msg;
    return rc;
}
}  // namespace

extern "C" {

const char* lph_last_error(void) { return g_err.c_str(); }

int lph_scene_load(const char* config_path, lph_scene** out)
{
    if (!config_path || !out) return herr(LP_EINVAL, "lph_scene_load: null argument");
    *out = nullptr;
    std::unique_ptr<lph_scene> s(new lph_scene);
    s->io.SetConfig(config_path);
    if (s->io.ParseFile()) return herr(LP_FAILURE, s->io.GetError());
    Emitter em;
    std::vector<float> p0, h, ang, rri;
    s->io.GetEmitterPointInPlane(p0);
    s->io.GetEmitterHeading(h);
    s->io.GetEmitterAngles(ang);
    s->io.GetEmitterRelativeRadiantIntensities(rri);
    em.SetNumberOfPhotons(s->io.GetNumberOfPhotons());
    em.SetPointInPlane(p0);
    em.SetDistanceToSurface(s->io.GetEmitterdistanceToSurface());
    em.SetRadiationCharacteristics(ang, rri);
    em.SetEmitterHeading(h);
    em.GetEmittermodel(s->pStart, s->vStart, s->vStartOcc);
    em.GetBasisVectors(s->eX, s->eY, s->eZ);
    *out = s.release();
    return LP_OK;
}

void lph_scene_free(lph_scene* s) { delete s; }

int lph_scene_sizes(const lph_scene* s, int64_t* ntri, int32_t* ntissue, int32_t* nstart)
{
    if (!s) return herr(LP_EINVAL, "null scene");
    std::vector<float> n;
    s->io.GetRefractionIndices(n);
    if (ntri) *ntri = s->io.GetMeshImage().size();
    if (ntissue) *ntissue = (int32_t)n.size();
    if (nstart) *nstart = (int32_t)s->vStartOcc.size();
    return LP_OK;
}

int lph_scene_mesh(const lph_scene* s, float* tri, float* nrm, int32_t* med)
{
    if (!s) return herr(LP_EINVAL, "null scene");
    const Mesh& m = s->io.GetMeshImage();
    if (tri) std::memcpy(tri, m.tri.data(), m.tri.size() * sizeof(float));
    if (nrm) std::memcpy(nrm, m.nrm.data(), m.nrm.size() * sizeof(float));
    if (med) std::memcpy(med, m.med.data(), m.med.size() * sizeof(int32_t));
    return LP_OK;
}

int lph_scene_materials(const lph_scene* s, float* nmuasg, int32_t* fetal_id)
{
    if (!s) return herr(LP_EINVAL, "null scene");
    std::vector<float> n, a, g, m;
    s->io.GetRefractionIndices(n);
    s->io.GetAbsorbtionCoefficients(a);
    s->io.GetAnisotropies(g);
    s->io.GetScatterCoefficients(m);
    if (nmuasg)
        for (size_t i = 0; i < n.size(); ++i) {
            nmuasg[4 * i] = n[i];
            nmuasg[4 * i + 1] = a[i];
            nmuasg[4 * i + 2] = m[i];
            nmuasg[4 * i + 3] = g[i];
        }
    if (fetal_id) *fetal_id = s->io.GetFetalID();
    return LP_OK;
}

int lph_scene_source(const lph_scene* s, float pstart[3], float* vstart, int32_t* vocc, float poi_basis[12])
{
    if (!s) return herr(LP_EINVAL, "null scene");
    std::vector<float> p0;
    s->io.GetEmitterPointInPlane(p0);
    if (pstart) std::memcpy(pstart, s->pStart.data(), 3 * sizeof(float));
    if (vstart) std::memcpy(vstart, s->vStart.data(), s->vStart.size() * sizeof(float));
    if (vocc)
        for (size_t i = 0; i < s->vStartOcc.size(); ++i) vocc[i] = s->vStartOcc[i];
    if (poi_basis)
        for (int k = 0; k < 3; ++k) {
            poi_basis[k] = p0[k];
            poi_basis[3 + k] = s->eX[k];
            poi_basis[6 + k] = s->eY[k];
            poi_basis[9 + k] = s->eZ[k];
        }
    return LP_OK;
}

int lph_scene_params(const lph_scene* s, float model_dims[3], float det_grid[6], float* factor, float* chance,
                     float* w_th, int64_t* nphotons)
{
    if (!s) return herr(LP_EINVAL, "null scene");
    std::vector<float> d, gr;
    s->io.GetDimensions(d);
    s->io.GetDetectionGrid(gr);
    if (model_dims) std::memcpy(model_dims, d.data(), 3 * sizeof(float));
    if (det_grid) std::memcpy(det_grid, gr.data(), 6 * sizeof(float));
    if (factor) *factor = s->io.GetPixelPerUnitFactor();
    if (chance) *chance = s->io.GetChance();
    if (w_th) *w_th = s->io.GetWeightThreshold();
    if (nphotons) *nphotons = s->io.GetNumberOfPhotons();
    return LP_OK;
}

int lph_scene_apply(const lph_scene* s, lp_context* ctx, uint64_t seed, int32_t rng_mode, int64_t max_steps)
{
    if (!s || !ctx) return herr(LP_EINVAL, "null argument");
    const Mesh& m = s->io.GetMeshImage();
    int32_t nt;
    lph_scene_sizes(s, nullptr, &nt, nullptr);
    std::vector<float> nm(4 * (size_t)nt), pb(12), d(3), gr(6);
    int32_t fid;
    float factor, chance, wth;
    lph_scene_materials(s, nm.data(), &fid);
    lph_scene_source(s, nullptr, nullptr, nullptr, pb.data());
    lph_scene_params(s, d.data(), gr.data(), &factor, &chance, &wth, nullptr);
    int rc;
    if ((rc = lp_set_mesh(ctx, m.tri.data(), m.nrm.data(), m.med.data(), m.size())) ||
        (rc = lp_set_materials(ctx, nm.data(), nt, fid)) ||
        (rc = lp_set_source(ctx, s->pStart.data(), s->vStart.data(), s->vStartOcc.data(),
                            (int32_t)s->vStartOcc.size(), pb.data())) ||
        (rc = lp_set_grid(ctx, d.data(), gr.data())) ||
        (rc = lp_set_params(ctx, chance, wth, factor, seed, rng_mode, max_steps)))
        return herr(rc, lp_last_error(ctx));
    return LP_OK;
}

int lph_emitter_model(const float p0[3], const float heading[3], float distance, const float* angles_deg,
                      const float* rri, int32_t nangles, int64_t nphotons, float pstart[3], float* vstart,
                      int32_t* vocc, float ex[3], float ey[3], float ez[3])
{
    if (!p0 || !heading || nangles < 1 || !angles_deg || !rri) return herr(LP_EINVAL, "lph_emitter_model: bad argument");
    Emitter em;
    em.SetNumberOfPhotons(nphotons);
    em.SetPointInPlane(std::vector<float>(p0, p0 + 3));
    em.SetDistanceToSurface(distance);
    em.SetRadiationCharacteristics(std::vector<float>(angles_deg, angles_deg + nangles),
                                   std::vector<float>(rri, rri + nangles));
    em.SetEmitterHeading(std::vector<float>(heading, heading + 3));
    std::vector<float> ps, vs, x, y, z;
    std::vector<int> oc;
    em.GetEmittermodel(ps, vs, oc);
    em.GetBasisVectors(x, y, z);
    if (pstart) std::memcpy(pstart, ps.data(), 3 * sizeof(float));
    if (vstart) std::memcpy(vstart, vs.data(), vs.size() * sizeof(float));
    if (vocc)
        for (size_t i = 0; i < oc.size(); ++i) vocc[i] = oc[i];
    if (ex) std::memcpy(ex, x.data(), 3 * sizeof(float));
    if (ey) std::memcpy(ey, y.data(), 3 * sizeof(float));
    if (ez) std::memcpy(ez, z.data(), 3 * sizeof(float));
    return LP_OK;
}

int lph_image(lp_context* ctx, int32_t which, int32_t component, int32_t extent[4], int32_t* ncomp, float* out)
{
    if (!ctx) return herr(LP_EINVAL, "null context");
    if (which != LP_TALLY_SFT && which != LP_TALLY_RTD && which != LP_TALLY_ARZ && which != LP_TALLY_PHIRZ)
        return herr(LP_EINVAL, "lph_image: SFT, RTD, ARZ or PHIRZ only");
    if (component < -1 || component > 2 || ((which == LP_TALLY_ARZ || which == LP_TALLY_PHIRZ) && component > 0))
        return herr(LP_EINVAL, "lph_image: component out of range");
    Image img;
    int rc = tally_image(ctx, which, component, img);
    if (rc) return herr(rc, lp_last_error(ctx));
    if (extent)
        for (int k = 0; k < 4; ++k) extent[k] = img.ext[k];
    if (ncomp) *ncomp = img.ncomp;
    if (out) std::memcpy(out, img.data.data(), img.data.size() * sizeof(float));
    return LP_OK;
}

float lph_scale_per_radius_1d(const float det_grid[6], int64_t nphotons, int32_t ir)
{
    Logs l;
    l.SetGridDimensions(std::vector<float>(det_grid, det_grid + 6));
    l.SetNumberOfPhotons(nphotons);
    return l.ScalePerRadius1D(ir);
}

float lph_scale_per_angle_1d(const float det_grid[6], int64_t nphotons, int32_t ia)
{
    Logs l;
    l.SetGridDimensions(std::vector<float>(det_grid, det_grid + 6));
    l.SetNumberOfPhotons(nphotons);
    return l.ScalePerAngle1D(ia);
}

float lph_scale_2d(const float det_grid[6], int64_t nphotons, int32_t ir, int32_t ia)
{
    Logs l;
    l.SetGridDimensions(std::vector<float>(det_grid, det_grid + 6));
    l.SetNumberOfPhotons(nphotons);
    return l.Scale2D(ir, ia);
}

int lph_profile(int32_t kind, const float* image, const int32_t extent[4], const float det_grid[6], int64_t nphotons,
                int32_t* rows, float* col0, float* col1)
{
    if (!image || !extent || !det_grid || !rows) return herr(LP_EINVAL, "lph_profile: null argument");
    Image im;
    for (int k = 0; k < 4; ++k) im.ext[k] = extent[k];
    if (im.nx() <= 0 || im.ny() <= 0) return herr(LP_EINVAL, "lph_profile: empty extent");
    im.data.assign(image, image + (size_t)im.nx() * im.ny());
    Logs l;
    l.SetGridDimensions(std::vector<float>(det_grid, det_grid + 6));
    l.SetNumberOfPhotons(nphotons);
    std::vector<float> a, b;
    if (kind == LPH_PER_RADIUS) l.ReflectancePerRadius(im, a, b);
    else if (kind == LPH_PER_ANGLE) l.ReflectancePerAngle(im, a, b);
    else if (kind == LPH_PER_Z) l.FluencePerZ(im, a, b);
    else return herr(LP_EINVAL, "lph_profile: unknown kind");
    *rows = (int32_t)a.size();
    if (col0) std::memcpy(col0, a.data(), a.size() * sizeof(float));
    if (col1) std::memcpy(col1, b.data(), b.size() * sizeof(float));
    return LP_OK;
}

int lph_surface_sums(const float* sft_image, const int32_t extent[4], float sums[3])
{
    if (!sft_image || !extent || !sums) return herr(LP_EINVAL, "lph_surface_sums: null argument");
    Image im;
    for (int k = 0; k < 4; ++k) im.ext[k] = extent[k];
    im.ncomp = 3;
    im.data.assign(sft_image, sft_image + (size_t)im.nx() * im.ny() * 3);
    Analysis a;
    a.SetRTdxyLogs(im);
    a.CalculateSurfaceSums();
    sums[0] = a.GetSurfRxySum();
    sums[1] = a.GetFetalRxySum();
    sums[2] = a.GetTransTxySum();
    return LP_OK;
}

int lph_run_config(const char* config_path, int64_t nphotons, uint64_t seed, int32_t device, const char* out_dir,
                   double sums[3])
{
    // src/main/main.cpp:101-320: import, emitter, propagation, analysis, logs
    ImportExport io;
    io.SetConfig(config_path ? config_path : "");
    if (io.ParseFile()) return herr(LP_FAILURE, io.GetError());
    std::vector<float> n, mua, g, mus, dims, grid, point, rri, angles, heading;
    io.GetRefractionIndices(n);
    io.GetAbsorbtionCoefficients(mua);
    io.GetAnisotropies(g);
    io.GetScatterCoefficients(mus);
    io.GetDimensions(dims);
    io.GetDetectionGrid(grid);
    io.GetEmitterAngles(angles);
    io.GetEmitterRelativeRadiantIntensities(rri);
    io.GetEmitterPointInPlane(point);
    io.GetEmitterHeading(heading);
    const int64_t N = nphotons > 0 ? nphotons : io.GetNumberOfPhotons();

    Emitter em;
    em.SetNumberOfPhotons(N);
    em.SetPointInPlane(point);
    em.SetDistanceToSurface(io.GetEmitterdistanceToSurface());
    em.SetRadiationCharacteristics(angles, rri);
    em.SetEmitterHeading(heading);
    std::vector<float> pStart, vStart, eX, eY, eZ;
    std::vector<int> vOcc;
    em.GetEmittermodel(pStart, vStart, vOcc);
    em.GetBasisVectors(eX, eY, eZ);

    CalculatePropagation cp;
    cp.SetDeviceType(device);
    cp.SetTissueProperties(n, mua, g, mus);
    cp.SetFetalTissueID(io.GetFetalID());
    cp.SetDimensions(dims);
    cp.SetDetectionGridDimensions(grid);
    cp.SetTriangles(io.GetMeshImage());
    cp.SetNormals(io.GetNormals());
    cp.SetInnerAndOuterTissueIndices(io.GetMediumIDs());
    cp.SetPixelPerUnitFactor(io.GetPixelPerUnitFactor());
    cp.SetStartTrajectories(vStart, vOcc);
    cp.SetStartPoint(pStart);
    cp.SetPOI(point);
    cp.SetBasisVectors(eX, eY, eZ);
    cp.SetChanceForRoulette(io.GetChance());
    cp.SetWeightThreshold(io.GetWeightThreshold());
    cp.SetNumberOfPhotons(N);
    cp.SetSeed(seed);
    if (cp.Init() || cp.Propagation()) return herr(LP_FAILURE, cp.GetError());

    Analysis an;
    an.SetNumberOfPhotons(N);
    an.SetRTdxyLogs(cp.GetSFTImage());
    an.CalculateSurfaceSums();
    an.PrintAnalysis();
    if (sums) {
        sums[0] = an.GetSurfRxySum() / (double)N;
        sums[1] = an.GetFetalRxySum() / (double)N;
        sums[2] = an.GetTransTxySum() / (double)N;
    }
    if (out_dir) {
        std::string dir = out_dir;
        if (!dir.empty() && dir.back() != '/') dir += '/';
        Logs lg;
        lg.SetSaveDir(dir);
        lg.SetDimensions(dims);
        lg.SetGridDimensions(grid);
        lg.SetNumberOfPhotons(N);
        std::string surf = io.GetSurfaceFilename().empty() ? "SurfRxy.txt" : io.GetSurfaceFilename();
        std::string trans = io.GetTransmissionFilename().empty() ? "TransRxy.txt" : io.GetTransmissionFilename();
        std::string fetal = io.GetFetalFilename().empty() ? "FetalRxy.txt" : io.GetFetalFilename();
        std::string ayz = io.GetRecordAyzFilename().empty() ? "FluencePrz.txt" : io.GetRecordAyzFilename();
        const int SURF = 0, FETAL = 1, TRANS = 2;  // Rxy_Log_* of src/IO/Logs.hpp:21-23
        int rc = lg.SaveSurfaceArr(surf, cp.GetSFTImage(SURF)) | lg.SaveRecordRd(surf, cp.GetRTdImage(SURF)) |
                 lg.CalcReflectancePerRadius(surf, cp.GetRTdImage(SURF)) |
                 lg.CalcReflectancePerAngle(surf, cp.GetRTdImage(SURF)) |
                 lg.CalcReflectancePerRadius(trans, cp.GetRTdImage(TRANS)) |
                 lg.CalcReflectancePerAngle(trans, cp.GetRTdImage(TRANS)) |
                 lg.CalcReflectancePerRadius(fetal, cp.GetRTdImage(FETAL)) |
                 lg.CalcReflectancePerAngle(fetal, cp.GetRTdImage(FETAL)) | lg.FluenceZ(ayz, cp.GetPhirzImage());
        if (rc) return herr(LP_FAILURE, "writing the log files in " + dir + " failed");
    }
    return LP_OK;
}

}  // extern "C"
