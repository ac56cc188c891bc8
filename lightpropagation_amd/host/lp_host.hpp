// lp_host.hpp -- host layer around liblp.so with the reference's class
// interfaces: ImportExport (configuration + OBJ + emitter file),
// Emitter, CalculatePropagation (the drop-in for the simulation call),
// Analysis and Logs (post-processing and text outputs).
//
// VTK and Boost are not used: a mesh is a plain Mesh, an image a plain
// Image with the reference's extents and float pixels, and the reference's
// regular expressions are matched by hand with the same whole-line rules.
// Method names, argument meaning and call order follow the reference
// (CalculatePropagation.hpp:212-389, ImportExport / Emitter / Analysis /
// Logs headers), so src/main/main.cpp:101-320 reads the same against these
// classes; what VTK drew (pictures, arrows) is not part of this layer.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lightprop.h"

namespace lph {

// Triangle mesh in the kernel's layout: vertex coordinates (9 per face),
// unit cell normals (3 per face), (inner, outer) tissue ids (2 per face).
struct Mesh {
    std::vector<float> tri;
    std::vector<float> nrm;
    std::vector<int32_t> med;
    int64_t size() const { return (int64_t)med.size() / 2; }
};

// A 2-D float image with VTK's extent convention (inclusive bounds, x
// fastest) and ncomp values per pixel.
struct Image {
    int ext[6] = {0, -1, 0, -1, 0, 0};
    int ncomp = 1;
    std::vector<float> data;
    void GetExtent(int e[6]) const;
    int nx() const { return ext[1] - ext[0] + 1; }
    int ny() const { return ext[3] - ext[2] + 1; }
    float GetScalarComponentAsFloat(int x, int y, int z, int c) const;
    void SetScalarComponentFromFloat(int x, int y, int z, int c, float v);
};

// ImportExport (src/IO/ImportExport.cpp) with ReadXML's parsing
// (src/IO/ReadXML.cpp).  Unset values keep ReadXML's defaults (chance 0.1,
// weight threshold 0.001, detection grid 1); PixelPerUnit defaults to 1
// rather than ReadXML's 0 (a zero factor makes every step length zero), and
// the fetal id to -1 (no tissue) where ReadXML leaves it uninitialised.
class ImportExport {
public:
    void SetVerbosity(bool v) { verbose = v; }
    void SetConfig(const std::string& filename) { configfile = filename; }
    // Returns 0 on success; the error text is in GetError().
    int ParseFile();
    const std::string& GetError() const { return error; }

    void GetRefractionIndices(std::vector<float>& v) const { v = refraction; }
    void GetAbsorbtionCoefficients(std::vector<float>& v) const { v = absorption; }
    void GetAnisotropies(std::vector<float>& v) const { v = anisotropy; }
    void GetScatterCoefficients(std::vector<float>& v) const { v = scatter; }
    void GetDimensions(std::vector<float>& v) const { v = dims; }
    void GetDetectionGrid(std::vector<float>& v) const { v = detGrid; }
    void GetEmitterAngles(std::vector<float>& v) const { v = angles; }
    void GetEmitterRelativeRadiantIntensities(std::vector<float>& v) const { v = rri; }
    void GetEmitterPointInPlane(std::vector<float>& v) const { v = emitterPoint; }
    void GetEmitterHeading(std::vector<float>& v) const { v = emitterHeading; }
    float GetEmitterdistanceToSurface() const { return emitterDistance; }
    const Mesh& GetMeshImage() const { return mesh; }
    const std::vector<float>& GetNormals() const { return mesh.nrm; }
    const std::vector<int32_t>& GetMediumIDs() const { return mesh.med; }
    void GetTissueNames(std::vector<std::string>& v) const { v = tissueNames; }
    void GetSurfaceIDVector(std::vector<int>& v) const { v = surfaceIDs; }
    void GetSurfaceNamesVector(std::vector<std::string>& v) const { v = surfaceIDnames; }
    int GetFetalID() const { return fetalID; }
    float GetPixelPerUnitFactor() const { return ppu; }
    float GetChance() const { return chance; }
    float GetWeightThreshold() const { return wThreshold; }
    int GetNumberOfPhotons() const { return numPhotons; }
    int GetNumberOfPhotonInteractions() const { return numPhotonInteractions; }
    bool GetSaveFile() const { return saveFile; }
    std::string GetSaveDir() const { return saveDir; }
    std::string GetSurfaceFilename() const { return surfaceFilename; }
    std::string GetFetalFilename() const { return fetalFilename; }
    std::string GetTransmissionFilename() const { return transmissionFilename; }
    std::string GetRecordAxzFilename() const { return recordAxzFilename; }
    std::string GetRecordAyzFilename() const { return recordAyzFilename; }
    std::string GetModelFilename() const { return modelfile; }

private:
    int ImportObjModel(const std::string& fname);
    int ImportEmitterModel(const std::string& fname);
    std::string Resolve(const std::string& path) const;
    int Fail(const std::string& msg);

    bool verbose = false;
    std::string configfile, error;
    std::vector<float> refraction, absorption, anisotropy, scatter;
    std::vector<std::string> tissueNames;
    std::vector<float> dims{1, 1, 1}, detGrid{1, 1, 1, 1, 1, 1};
    std::vector<float> emitterPoint{0, 0, 0}, emitterHeading{0, 0, 1};
    float emitterDistance = 0;
    std::vector<float> angles, rri;
    Mesh mesh;
    std::vector<int> surfaceIDs;
    std::vector<std::string> surfaceIDnames;
    int fetalID = -1;
    float ppu = 1, chance = 0.1f, wThreshold = 0.001f;
    int numPhotons = 0, numPhotonInteractions = 0;
    bool saveFile = false;
    std::string modelfile, emitterFile, saveDir;
    std::string surfaceFilename, fetalFilename, transmissionFilename, recordAxzFilename, recordAyzFilename;
};

// Emitter (src/Model/Emitter.cpp:186-339).
class Emitter {
public:
    void SetVerbosity(bool v) { verbose = v; }
    void SetNumberOfPhotons(int64_t n) { numberOfPhotons = n; }
    void SetPointInPlane(const std::vector<float>& p) { p0 = p; }
    void SetDistanceToSurface(float d) { distance = d; }
    void SetRadiationCharacteristics(const std::vector<float>& a, const std::vector<float>& r)
    {
        angles = a;
        rri = r;
    }
    void SetEmitterHeading(const std::vector<float>& h) { heading = h; }
    // start point, directions (3 per angle) and photons per direction
    void GetEmittermodel(std::vector<float>& pStart, std::vector<float>& vStart, std::vector<int>& vStartOcc);
    void GetBasisVectors(std::vector<float>& eX, std::vector<float>& eY, std::vector<float>& eZ);

private:
    void Calculate();
    bool verbose = false;
    int64_t numberOfPhotons = 0;
    std::vector<float> p0{0, 0, 0}, heading{0, 0, 1}, angles{0}, rri{1};
    float distance = 0;
    std::vector<float> position, startVector, ex, ey, ez;
    std::vector<int> occ;
};

// CalculatePropagation (src/Simulation/CalculatePropagation.hpp:212-389):
// the same setters and call order, forwarding to the C ABI.  Init() builds
// the transport context; Propagation() runs the photons in one persistent
// launch; the Get*Image calls return the tallies as the reference's images.
class CalculatePropagation {
public:
    CalculatePropagation() = default;
    ~CalculatePropagation();
    CalculatePropagation(const CalculatePropagation&) = delete;
    CalculatePropagation& operator=(const CalculatePropagation&) = delete;

    void SetVerbosity(bool v) { verbose = v; }
    void SetShowPercentDone(bool) {}
    void SetDeviceType(int device) { this->device = device; }
    // argument order (n, mua, g, mus) of CalculatePropagation.hpp:333
    void SetTissueProperties(const std::vector<float>& refractionInd, const std::vector<float>& absorbtionCoeffs,
                             const std::vector<float>& anisotropies, const std::vector<float>& scatterCoeffs);
    void SetFetalTissueID(int id) { fetalID = id; }
    void SetDimensions(const std::vector<float>& d) { dims = d; }
    void SetDetectionGridDimensions(const std::vector<float>& g) { detGrid = g; }
    void SetDetectionGridDimensions(float x, float y, float z, float r, float phi, float rho)
    {
        detGrid = {x, y, z, r, phi, rho};
    }
    void SetTriangles(const Mesh& m) { mesh.tri = m.tri; }
    void SetNormals(const std::vector<float>& n) { mesh.nrm = n; }
    void SetInnerAndOuterTissueIndices(const std::vector<int32_t>& med) { mesh.med = med; }
    void SetPixelPerUnitFactor(float f) { factor = f; }
    void SetStartTrajectories(const std::vector<float>& v, const std::vector<int>& occ)
    {
        vStart = v;
        vStartOcc = occ;
    }
    void SetStartTrajectory(float x, float y, float z)
    {
        vStart = {x, y, z};
        vStartOcc = {1};
    }
    void SetStartPoint(const std::vector<float>& p) { pStart = p; }
    void SetStartPoint(float x, float y, float z) { pStart = {x, y, z}; }
    void SetPOI(const std::vector<float>& p) { poi = p; }
    void SetPOI(float x, float y, float z) { poi = {x, y, z}; }
    void SetBasisVectors(const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& z)
    {
        eX = x;
        eY = y;
        eZ = z;
    }
    void SetChanceForRoulette(float c) { chance = c; }
    void SetWeightThreshold(float w) { wThreshold = w; }
    void SetNumberOfPhotons(int64_t n) { numberOfPhotons = n; }
    void SetNumberOfPhotonInteractions(int) {}  // unused by the reference as well
    // seed of the per-photon streams (the reference seeds from the clock)
    void SetSeed(uint64_t s) { seed = s; }
    void SetNumberGenerator(int rngMode) { rng = rngMode; }

    int Init();         // 0 on success; error text in GetError()
    int Propagation();  // runs GetNumberOfPhotons() photons
    void PrintTimeSpent() const;
    const std::string& GetError() const { return error; }
    int64_t GetNumberOfPhotons() const { return numberOfPhotons; }
    lp_context* GetContext() { return ctx; }

    Image GetSFTImage() const;       // 3 components: R, F, T
    Image GetSFTImage(int c) const;  // one component
    Image GetRTdImage() const;
    Image GetRTdImage(int c) const;
    Image GetArzImage() const;    // absorbed weight over (r, z)
    Image GetPhirzImage() const;  // absorbed weight / mua over (r, z)

private:
    int Fail(const std::string& msg);
    bool verbose = false;
    int device = 0;
    std::vector<float> nmuasg;
    int fetalID = -1;
    std::vector<float> dims, detGrid;
    Mesh mesh;
    float factor = 1, chance = 0.1f, wThreshold = 1e-4f;
    std::vector<float> vStart{0, 0, 1}, pStart{0, 0, 0}, poi{0, 0, 0}, eX{1, 0, 0}, eY{0, 1, 0}, eZ{0, 0, 1};
    std::vector<int> vStartOcc{1};
    int64_t numberOfPhotons = 0;
    uint64_t seed = 1;
    int rng = LP_RNG_PHILOX;
    lp_context* ctx = nullptr;
    double kernelMs = 0;
    std::string error;
};

// Analysis (src/Analysis/Analysis.cpp:144-156, 447-463, 529-538).
class Analysis {
public:
    void SetVerbosity(bool) {}
    void SetNumberOfPhotons(int64_t n) { numberPhotons = n; }
    void SetRTdxyLogs(const Image& sft);  // the 3-component SFT image
    void CalculateSurfaceSums();
    void PrintAnalysis() const;
    float GetSurfRxySum() const { return surfRxySum; }
    float GetFetalRxySum() const { return fetalRxySum; }
    float GetTransTxySum() const { return transTxySum; }

private:
    int64_t numberPhotons = 0;
    std::vector<float> surfRxyLog, fetalRxyLog, transTxyLog;
    float surfRxySum = 0, fetalRxySum = 0, transTxySum = 0;
};

// Logs: the text outputs and MCML normalisations of src/IO/Logs.cpp.
class Logs {
public:
    void SetSaveDir(const std::string& d) { saveDir = d; }
    void SetDimensions(const std::vector<float>& d) { dims = d; }
    void SetGridDimensions(const std::vector<float>& g) { gridDims = g; }
    void SetNumberOfPhotons(int64_t n) { numberPhotons = n; }
    void SetPixelPerUnitFactor(float) {}

    float ScalePerRadius1D(int ir) const;
    float ScalePerAngle1D(int ia) const;
    float Scale2D(int ir, int ia) const;

    // Profiles as (column 0, column 1) rows, exactly what the files hold.
    void ReflectancePerRadius(const Image& rtd, std::vector<float>& c0, std::vector<float>& c1) const;
    void ReflectancePerAngle(const Image& rtd, std::vector<float>& c0, std::vector<float>& c1) const;
    void FluencePerZ(const Image& phirz, std::vector<float>& c0, std::vector<float>& c1) const;

    // The files (saveDir + name); each returns 0 on success.
    int SaveSurfaceArr(const std::string& fname, const Image& image) const;
    int SaveRecordRd(const std::string& fname, const Image& rtd) const;
    int CalcReflectancePerRadius(const std::string& fname, const Image& rtd) const;
    int CalcReflectancePerAngle(const std::string& fname, const Image& rtd) const;
    int FluenceZ(const std::string& fname, const Image& phirz) const;

private:
    std::string saveDir;
    std::vector<float> dims{1, 1, 1}, gridDims{1, 1, 1, 1, 1, 1};
    int64_t numberPhotons = 0;
};

}  // namespace lph
