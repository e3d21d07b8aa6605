// HDF5 / XDMF output of structured fields (C ABI in include/rf_h5.h).
//
// One rank's share of write_Simple_pHDF5_Str, write_HDF5_attributes and
// write_XMF_Elements (writeResultFile_RF.f90): the same dataset name, rank,
// dimension order, element and attribute types, and XDMF text, so files from
// this library read like the reference's in its post-processing.  The HDF5
// C API is used directly (serial driver: one GPU, one writer).
#include "../../include/rf_h5.h"

#include <hdf5.h>
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// mkdir -p for the h5/ and xmf/ folders
bool make_dirs(const std::string& path) {
    if (path.empty()) return true;
    std::string cur;
    size_t i = 0;
    while (i <= path.size()) {
        const size_t j = path.find('/', i);
        cur = path.substr(0, j == std::string::npos ? path.size() : j);
        if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
        if (j == std::string::npos) break;
        i = j + 1;
    }
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string join(const std::string& a, const std::string& b) {
    if (a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

void copy_out(const std::string& s, char* dst, int32_t cap) {
    if (!dst || cap <= 0) return;
    std::snprintf(dst, (size_t)cap, "%s", s.c_str());
}

// Fortran F25.10 edit descriptor for finite values
std::string f25_10(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%25.10f", v);
    return b;
}

// RAII holder for an HDF5 identifier and its close function
struct H5Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Id() {
        if (id >= 0 && close) close(id);
    }
    H5Id(const H5Id&) = delete;
    H5Id& operator=(const H5Id&) = delete;
};

int attr_write(hid_t file, const char* name, hid_t ftype, hid_t mtype, int rank, hsize_t n, const void* data) {
    H5Id space(rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &n, &n), H5Sclose);
    if (space.id < 0) return fail(RF_H5_E_IO, std::string("dataspace for attribute ") + name);
    if (H5Aexists(file, name) > 0 && H5Adelete(file, name) < 0)
        return fail(RF_H5_E_IO, std::string("cannot replace attribute ") + name);
    H5Id at(H5Acreate2(file, name, ftype, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (at.id < 0) return fail(RF_H5_E_IO, std::string("cannot create attribute ") + name);
    if (H5Awrite(at.id, mtype, data) < 0) return fail(RF_H5_E_IO, std::string("cannot write attribute ") + name);
    return RF_H5_OK;
}

}  // namespace

extern "C" const char* rf_h5_last_error(void) { return g_err.c_str(); }

extern "C" int rf_h5_write_xmf(const char* xmfPath, const char* h5FileName, const char* h5ToXmfPath,
                               int32_t nDim, const double* xMin, const double* xMax, const double* xStep,
                               const char* dsetName) {
    g_err.clear();
    if (!xmfPath || !h5FileName || !h5ToXmfPath || !xMin || !xMax || !xStep || !dsetName)
        return fail(RF_H5_E_ARG, "NULL argument");
    if (nDim != 2 && nDim != 3) return fail(RF_H5_E_ARG, "the XDMF model exists for nDim 2 and 3");
    // total_xNStep = nint((xMax - xMin)/xStep) + 1, listed last axis first
    std::string dimText;
    for (int i = nDim - 1; i >= 0; i--) {
        const long long n = std::llround((xMax[i] - xMin[i]) / xStep[i]) + 1;
        if (!dimText.empty()) dimText += " ";
        dimText += std::to_string(n);
    }
    FILE* f = std::fopen(xmfPath, "w");
    if (!f) return fail(RF_H5_E_IO, std::string("cannot open ") + xmfPath);
    const char* topo = nDim == 2 ? "2DCoRectMesh" : "3DCoRectMesh";
    const char* geo = nDim == 2 ? "ORIGIN_DXDY" : "ORIGIN_DXDYDZ";
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n<Xdmf Version=\"2.0\">\n <Domain>\n");
    std::fprintf(f, "   <DataItem Name=\"RF_1\" Format=\"HDF\" DataType=\"Float\" Precision=\"8\" Dimensions=\"%s\">\n",
                 dimText.c_str());
    std::fprintf(f, "        %s/%s:/%s\n", h5ToXmfPath, h5FileName, dsetName);
    std::fprintf(f, "   </DataItem>\n  <Grid GridType=\"Collection\" CollectionType=\"Spatial\">\n");
    std::fprintf(f, "   <Grid Name=\"Group1\">\n");
    std::fprintf(f, "     <Topology TopologyType=\"%s\" Dimensions=\"%s\"/>\n", topo, dimText.c_str());
    std::fprintf(f, "     <Geometry GeometryType=\"%s\">\n", geo);
    // '(A,F25.10)' over (' ', v_d, ' ', v_{d-1}, ...): format reversion puts
    // each value on its own record, the first after four blanks
    auto values = [&](const char* item, const double* v) {
        std::fprintf(f, "   <DataItem Name=\"%s\" Format=\"XML\" DataType=\"Float\" Precision=\"8\" Dimensions=\"%d\">\n",
                     item, nDim);
        for (int i = nDim - 1; i >= 0; i--)
            std::fprintf(f, "%s%s\n", i == nDim - 1 ? "    " : " ", f25_10(v[i]).c_str());
        std::fprintf(f, "   </DataItem>\n");
    };
    values("origin", xMin);
    values("step", xStep);
    std::fprintf(f, "     </Geometry>\n");
    std::fprintf(f, "     <Attribute Name=\"RF1\" Center=\"Node\" AttributeType=\"Scalar\">\n");
    std::fprintf(f, "       <DataItem Reference=\"XML\">\n");
    std::fprintf(f, "         /Xdmf/Domain/DataItem[@Name=\"RF_1\"]\n");
    std::fprintf(f, "       </DataItem>\n     </Attribute>\n");
    std::fprintf(f, "   </Grid>\n  </Grid>\n </Domain>\n</Xdmf>\n");
    if (std::fclose(f) != 0) return fail(RF_H5_E_IO, std::string("cannot write ") + xmfPath);
    return RF_H5_OK;
}

extern "C" int rf_h5_write_simple(const char* folder, const char* name, int32_t nDim,
                                  const int32_t* minPos, const int32_t* maxPos, const int32_t* origin,
                                  const int32_t* xNStep, const double* xStep, const double* xMinGlob,
                                  const int32_t* xNStepGlob, const double* field,
                                  char* h5Path, char* xmfPath, int32_t pathCap) {
    g_err.clear();
    if (!folder || !name || !minPos || !maxPos || !origin || !xNStep || !xStep || !xMinGlob || !xNStepGlob ||
        !field)
        return fail(RF_H5_E_ARG, "NULL argument");
    if (nDim < 1 || nDim > 3) return fail(RF_H5_E_ARG, "nDim must be 1, 2 or 3");
    hsize_t dims[3], offset[3], count[3];
    int64_t cnt[3];
    for (int i = 0; i < nDim; i++) {
        if (xNStep[i] < 1 || xNStepGlob[i] < 1) return fail(RF_H5_E_ARG, "grid counts must be >= 1");
        if (minPos[i] < 1 || maxPos[i] < minPos[i] || maxPos[i] > xNStep[i])
            return fail(RF_H5_E_ARG, "window minPos..maxPos outside the field");
        cnt[i] = (int64_t)maxPos[i] - minPos[i] + 1;
        if (origin[i] < 1 || origin[i] - 1 + cnt[i] > xNStepGlob[i])
            return fail(RF_H5_E_ARG, "window outside the global grid");
        // HDF5 (C order) lists the Fortran dimensions last first
        const int c = nDim - 1 - i;
        dims[c] = (hsize_t)xNStepGlob[i];
        offset[c] = (hsize_t)(origin[i] - 1);
        count[c] = (hsize_t)cnt[i];
    }
    const std::string fileH5 = std::string(name) + ".h5";
    const std::string h5Dir = join(folder, "h5"), xmfDir = join(folder, "xmf");
    const std::string full = join(h5Dir, fileH5);
    const std::string xmf = join(xmfDir, std::string(name) + ".xmf");
    copy_out(full, h5Path, pathCap);
    copy_out(xmf, xmfPath, pathCap);
    if (!make_dirs(h5Dir)) return fail(RF_H5_E_IO, "cannot create " + h5Dir);

    // the window, packed x fastest
    int64_t n = 1;
    for (int i = 0; i < nDim; i++) n *= cnt[i];
    std::vector<double> buf((size_t)n);
    const int64_t s1 = xNStep[0], s2 = nDim > 1 ? (int64_t)xNStep[0] * xNStep[1] : 0;
    const int64_t c1 = nDim > 1 ? cnt[1] : 1, c2 = nDim > 2 ? cnt[2] : 1;
    size_t o = 0;
    for (int64_t k = 0; k < c2; k++)
        for (int64_t j = 0; j < c1; j++) {
            const int64_t base = (minPos[0] - 1) + (nDim > 1 ? (minPos[1] - 1 + j) * s1 : 0) +
                                 (nDim > 2 ? (minPos[2] - 1 + k) * s2 : 0);
            std::memcpy(&buf[o], field + base, (size_t)cnt[0] * sizeof(double));
            o += (size_t)cnt[0];
        }

    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Id file(H5Fcreate(full.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(RF_H5_E_IO, "cannot create " + full);
    H5Id fspace(H5Screate_simple(nDim, dims, nullptr), H5Sclose);
    if (fspace.id < 0) return fail(RF_H5_E_IO, "dataspace of samples");
    H5Id dset(H5Dcreate2(file.id, "samples", H5T_NATIVE_DOUBLE, fspace.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
              H5Dclose);
    if (dset.id < 0) return fail(RF_H5_E_IO, "cannot create dataset samples");
    H5Id mspace(H5Screate_simple(nDim, count, nullptr), H5Sclose);
    if (mspace.id < 0 || H5Sselect_hyperslab(fspace.id, H5S_SELECT_SET, offset, nullptr, count, nullptr) < 0)
        return fail(RF_H5_E_IO, "hyperslab of samples");
    if (H5Dwrite(dset.id, H5T_NATIVE_DOUBLE, mspace.id, fspace.id, H5P_DEFAULT, buf.data()) < 0)
        return fail(RF_H5_E_IO, "cannot write samples");
    if (nDim < 2) return RF_H5_OK;              // no XDMF model in 1-D
    if (!make_dirs(xmfDir)) return fail(RF_H5_E_IO, "cannot create " + xmfDir);
    double xMax[3];
    for (int i = 0; i < nDim; i++) xMax[i] = xMinGlob[i] + (double)(xNStepGlob[i] - 1) * xStep[i];
    return rf_h5_write_xmf(xmf.c_str(), fileH5.c_str(), "../h5", nDim, xMinGlob, xMax, xStep, "samples");
}

extern "C" int rf_h5_write_attributes(const char* h5Path, const rf_h5_attrs* a) {
    g_err.clear();
    if (!h5Path || !a) return fail(RF_H5_E_ARG, "NULL argument");
    if (a->nDim < 1 || a->nDim > 3) return fail(RF_H5_E_ARG, "nDim must be 1, 2 or 3");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Id file(H5Fopen(h5Path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(RF_H5_E_IO, std::string("cannot open ") + h5Path);
    const hid_t fid = file.id;
    const hsize_t d = (hsize_t)a->nDim;
    const struct { const char* n; const int32_t* v; } ints[] = {
        {"nb_procs", &a->nb_procs}, {"nDim", &a->nDim}, {"Nmc", &a->Nmc}, {"method", &a->method},
        {"seedStart", &a->seedStart}, {"corrMod", &a->corrMod}, {"margiFirst", &a->margiFirst},
        {"localizationLevel", &a->localizationLevel}};
    for (const auto& e : ints) {
        const int rc = attr_write(fid, e.n, H5T_STD_I32LE, H5T_NATIVE_INT32, 0, 1, e.v);
        if (rc) return rc;
    }
    const struct { const char* n; const int32_t* v; } ivec[] = {{"kNStep", a->kNStep}, {"nFields", a->nFields}};
    for (const auto& e : ivec) {
        const int rc = attr_write(fid, e.n, H5T_STD_I32LE, H5T_NATIVE_INT32, 1, d, e.v);
        if (rc) return rc;
    }
    const struct { const char* n; const double* v; } rvec[] = {
        {"xMinGlob", a->xMinGlob}, {"xMaxGlob", a->xMaxGlob}, {"xStep", a->xStep}, {"corrL", a->corrL},
        {"overlap", a->overlap}, {"procExtent", a->procExtent}, {"kMax_out", a->kMax_out}};
    for (const auto& e : rvec) {
        const int rc = attr_write(fid, e.n, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 1, d, e.v);
        if (rc) return rc;
    }
    return RF_H5_OK;
}

namespace {

int attr_read(hid_t file, const char* name, hid_t mtype, size_t n, void* data) {
    if (H5Aexists(file, name) <= 0) return fail(RF_H5_E_IO, std::string("attribute ") + name + " missing");
    H5Id at(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    if (at.id < 0) return fail(RF_H5_E_IO, std::string("cannot open attribute ") + name);
    H5Id sp(H5Aget_space(at.id), H5Sclose);
    const hssize_t cnt = sp.id < 0 ? -1 : H5Sget_simple_extent_npoints(sp.id);
    if (cnt < 0 || (size_t)cnt < n) return fail(RF_H5_E_IO, std::string("attribute ") + name + " too short");
    std::vector<char> buf((size_t)cnt * H5Tget_size(mtype));
    if (H5Aread(at.id, mtype, buf.data()) < 0) return fail(RF_H5_E_IO, std::string("cannot read attribute ") + name);
    std::memcpy(data, buf.data(), n * H5Tget_size(mtype));
    return RF_H5_OK;
}

}  // namespace

extern "C" int rf_h5_read_box_info(const char* h5Path, int32_t nDim, const double* xMinLoc, const double* xMaxLoc,
                                   rf_h5_box* box) {
    g_err.clear();
    if (!h5Path || !xMinLoc || !xMaxLoc || !box) return fail(RF_H5_E_ARG, "NULL argument");
    if (nDim < 1 || nDim > 3) return fail(RF_H5_E_ARG, "nDim must be 1, 2 or 3");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Id file(H5Fopen(h5Path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(RF_H5_E_IO, std::string("cannot open ") + h5Path);
    rf_h5_box b{};
    int rc = attr_read(file.id, "nDim", H5T_NATIVE_INT32, 1, &b.nDim);
    if (!rc) rc = attr_read(file.id, "Nmc", H5T_NATIVE_INT32, 1, &b.Nmc);
    if (rc) return rc;
    if (b.nDim != nDim) return fail(RF_H5_E_ARG, "nDim differs from the file's attribute");
    rc = attr_read(file.id, "xMinGlob", H5T_NATIVE_DOUBLE, (size_t)nDim, b.xMinGlob);
    if (!rc) rc = attr_read(file.id, "xMaxGlob", H5T_NATIVE_DOUBLE, (size_t)nDim, b.xMaxGlob);
    if (!rc) rc = attr_read(file.id, "xStep", H5T_NATIVE_DOUBLE, (size_t)nDim, b.xStep);
    if (rc) return rc;
    for (int i = 0; i < nDim; i++) {
        if (!(b.xStep[i] > 0.0)) return fail(RF_H5_E_IO, "xStep attribute must be > 0");
        b.xNStep[i] = (int32_t)std::llround((b.xMaxGlob[i] - b.xMinGlob[i]) / b.xStep[i]) + 1;
        int64_t mn = (int64_t)std::floor((xMinLoc[i] - b.xMinGlob[i]) / b.xStep[i]) + 1;
        int64_t mx = (int64_t)std::ceil((xMaxLoc[i] - b.xMinGlob[i]) / b.xStep[i]) + 1;
        if (mn < 1) mn = 1;
        if (mx > b.xNStep[i]) mx = b.xNStep[i];
        if (mx < mn) return fail(RF_H5_E_ARG, "the nodes lie outside the grid");
        b.minPos[i] = (int32_t)mn;
        b.maxPos[i] = (int32_t)mx;
    }
    *box = b;
    return RF_H5_OK;
}

extern "C" int rf_h5_read_box(const char* h5Path, const rf_h5_box* box, double* out) {
    g_err.clear();
    if (!h5Path || !box || !out) return fail(RF_H5_E_ARG, "NULL argument");
    const int nDim = box->nDim;
    if (nDim < 1 || nDim > 3) return fail(RF_H5_E_ARG, "nDim must be 1, 2 or 3");
    hsize_t offset[3], count[3];
    for (int i = 0; i < nDim; i++) {
        if (box->minPos[i] < 1 || box->maxPos[i] < box->minPos[i]) return fail(RF_H5_E_ARG, "bad box");
        offset[nDim - 1 - i] = (hsize_t)(box->minPos[i] - 1);
        count[nDim - 1 - i] = (hsize_t)(box->maxPos[i] - box->minPos[i] + 1);
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Id file(H5Fopen(h5Path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(RF_H5_E_IO, std::string("cannot open ") + h5Path);
    H5Id dset(H5Dopen2(file.id, "samples", H5P_DEFAULT), H5Dclose);
    if (dset.id < 0) return fail(RF_H5_E_IO, "dataset samples missing");
    H5Id fspace(H5Dget_space(dset.id), H5Sclose);
    if (fspace.id < 0 || H5Sget_simple_extent_ndims(fspace.id) != nDim)
        return fail(RF_H5_E_IO, "samples has another rank");
    hsize_t dims[3];
    H5Sget_simple_extent_dims(fspace.id, dims, nullptr);
    for (int c = 0; c < nDim; c++)
        if (offset[c] + count[c] > dims[c]) return fail(RF_H5_E_ARG, "box outside the samples dataset");
    H5Id mspace(H5Screate_simple(nDim, count, nullptr), H5Sclose);
    if (mspace.id < 0 || H5Sselect_hyperslab(fspace.id, H5S_SELECT_SET, offset, nullptr, count, nullptr) < 0)
        return fail(RF_H5_E_IO, "hyperslab of samples");
    if (H5Dread(dset.id, H5T_NATIVE_DOUBLE, mspace.id, fspace.id, H5P_DEFAULT, out) < 0)
        return fail(RF_H5_E_IO, "cannot read samples");
    return RF_H5_OK;
}
