// readUNV (readUNV_RF.f90:14-215 with prepareCoordinates :519-572,
// prepareConnectivity :578-676, readCoordinates :689-776, readConnectivity
// :782-915) on one rank: node coordinates (dataset 2411 / 781) and element
// connectivity (2412 / 780) of an I-DEAS universal file, the input of
// interpolateToMesh.  Host code; the coordinates go to the device for the
// interpolation kernel.
//
// Records are read as the reference's list-directed reads see them: a
// record's "tag" is its first blank- or comma-delimited token (blank records
// are skipped), numbers accept the D exponent.  Every record whose tag is
// 2411/781 or 2412/780 outside a dataset opens one.  The reference counts only
// FE descriptor 115 (8-node brick) in its dimensioning pass and leaves the
// records of any other element type to be misread; such files are refused.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "rf_internal.h"

struct rf_unv_mesh {
    int nDim = 0;
    std::vector<double> coords;      // nDim x nNodes
    std::vector<int32_t> connect;    // maxConnect x nElem, -1 padded
    std::vector<int32_t> sizes;      // nodes per element
    int64_t nNodes = 0, nElem = 0;
    int32_t maxConnect = 0;
    bool monoType = true;
};

namespace {

struct Records {
    std::ifstream in;
    std::string line;
    int64_t lineNb = 0;
    // next non-blank record; false at end of file
    bool next() {
        while (std::getline(in, line)) {
            lineNb++;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.find_first_not_of(" \t") != std::string::npos) return true;
        }
        return false;
    }
    std::string tag() const {
        const size_t b = line.find_first_not_of(" \t");
        const size_t e = line.find_first_of(" \t,/", b);
        return line.substr(b, e == std::string::npos ? std::string::npos : e - b);
    }
};

// the numbers of one record (blank or comma separated, D exponent allowed)
bool numbers(const std::string& rec, std::vector<double>* v) {
    v->clear();
    std::string tok;
    auto flush = [&]() -> bool {
        if (tok.empty()) return true;
        for (char& c : tok)
            if (c == 'D' || c == 'd') c = 'E';
        char* end = nullptr;
        const double x = std::strtod(tok.c_str(), &end);
        if (!end || *end != '\0') return false;
        v->push_back(x);
        tok.clear();
        return true;
    };
    for (char c : rec) {
        if (c == ' ' || c == '\t' || c == ',') {
            if (!flush()) return false;
        } else {
            tok.push_back(c);
        }
    }
    return flush();
}

int err_at(const Records& r, const std::string& what) {
    return rf_set_error(nullptr, RF_E_ARG, "UNV line " + std::to_string(r.lineNb) + ": " + what);
}

}  // namespace

extern "C" {

int rf_unv_read(const char* path, int32_t nDim, rf_unv_mesh** out) {
    if (!out) return rf_set_error(nullptr, RF_E_ARG, "mesh out-pointer is NULL");
    *out = nullptr;
    if (!path) return rf_set_error(nullptr, RF_E_ARG, "path is NULL");
    if (nDim < 1 || nDim > 3) return rf_set_error(nullptr, RF_E_ARG, "nDim must be 1, 2 or 3");
    Records r;
    r.in.open(path);
    if (!r.in) return rf_set_error(nullptr, RF_E_ARG, std::string("cannot open ") + path);
    std::unique_ptr<rf_unv_mesh> M(new (std::nothrow) rf_unv_mesh());
    if (!M) return rf_set_error(nullptr, RF_E_NOMEM, "mesh allocation failed");
    M->nDim = nDim;
    std::vector<double> v;
    std::vector<std::vector<int32_t>> elems;
    while (r.next()) {
        const std::string t = r.tag();
        if (t == "2411" || t == "781") {
            // node records: "label csys dsys color", then the coordinates
            while (true) {
                if (!r.next()) return err_at(r, "end of file inside dataset 2411");
                if (r.tag() == "-1") break;
                if (!r.next()) return err_at(r, "node without coordinates");
                if (!numbers(r.line, &v) || (int)v.size() < nDim) return err_at(r, "bad coordinate record");
                for (int a = 0; a < nDim; a++) M->coords.push_back(v[(size_t)a]);
                M->nNodes++;
            }
        } else if (t == "2412" || t == "780") {
            // element records: "label FEid phys mat color nNodes", then the
            // node labels, eight per record
            while (true) {
                if (!r.next()) return err_at(r, "end of file inside dataset 2412");
                if (r.tag() == "-1") break;
                if (!numbers(r.line, &v) || v.size() < 6) return err_at(r, "bad element header");
                const int fe = (int)v[1], nn = (int)v[5];
                if (fe != 115)
                    return rf_set_error(nullptr, RF_E_UNSUPPORTED,
                                        "UNV line " + std::to_string(r.lineNb) + ": element type " + std::to_string(fe) +
                                            " (the reference reads FE descriptor 115 only)");
                if (nn < 1 || nn > 50) return err_at(r, "bad node count");
                std::vector<int32_t> e;
                for (int i = 0; i < (nn + 7) / 8; i++) {
                    if (!r.next() || !numbers(r.line, &v)) return err_at(r, "bad connectivity record");
                    const int want = std::min(8, nn - 8 * i);
                    if ((int)v.size() < want) return err_at(r, "short connectivity record");
                    for (int k = 0; k < want; k++) e.push_back((int32_t)v[(size_t)k]);
                }
                if (M->maxConnect == 0) M->maxConnect = nn;
                if (M->maxConnect != nn) {
                    if (M->maxConnect < nn) M->maxConnect = nn;
                    M->monoType = false;
                }
                elems.push_back(std::move(e));
            }
        }
    }
    M->nElem = (int64_t)elems.size();
    // readUNV_RF.f90:77: too few nodes or elements for the ranks (one here)
    if (M->nNodes < 1 || M->nElem < 1) return rf_set_error(nullptr, RF_E_ARG, "UNV file without nodes or elements");
    M->connect.assign((size_t)(M->maxConnect * M->nElem), -1);
    M->sizes.resize((size_t)M->nElem);
    for (int64_t i = 0; i < M->nElem; i++) {
        const auto& e = elems[(size_t)i];
        std::memcpy(&M->connect[(size_t)(i * M->maxConnect)], e.data(), e.size() * sizeof(int32_t));
        // monoType: every entry of sizeList is size(connectList, 1)
        M->sizes[(size_t)i] = M->monoType ? M->maxConnect : (int32_t)e.size();
    }
    *out = M.release();
    return RF_OK;
}

int rf_unv_sizes(const rf_unv_mesh* m, int64_t* nNodes, int64_t* nElem, int32_t* maxConnect, int32_t* monoType) {
    if (!m) return rf_set_error(nullptr, RF_E_ARG, "mesh is NULL");
    if (nNodes) *nNodes = m->nNodes;
    if (nElem) *nElem = m->nElem;
    if (maxConnect) *maxConnect = m->maxConnect;
    if (monoType) *monoType = m->monoType ? 1 : 0;
    return RF_OK;
}

int rf_unv_copy(const rf_unv_mesh* m, double* coords, int32_t* connect, int32_t* sizeList) {
    if (!m) return rf_set_error(nullptr, RF_E_ARG, "mesh is NULL");
    if (coords) std::memcpy(coords, m->coords.data(), m->coords.size() * sizeof(double));
    if (connect) std::memcpy(connect, m->connect.data(), m->connect.size() * sizeof(int32_t));
    if (sizeList) std::memcpy(sizeList, m->sizes.data(), m->sizes.size() * sizeof(int32_t));
    return RF_OK;
}

int rf_unv_destroy(rf_unv_mesh* m) {
    delete m;
    return RF_OK;
}

}  // extern "C"
