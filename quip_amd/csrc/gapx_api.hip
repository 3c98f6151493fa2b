// Host side of libgapx.so: the C ABI of include/gapx.h, the GAP XML model
// reader, device-resident model state and the per-call pipeline
//   neighbour list -> SOAP forward -> GEMM1 (K, fused GP epilogue) ->
//   GEMM2 (G) -> SOAP adjoint -> distance_2b -> deterministic gather.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gapx.h"
#include "gapx_kernels.h"

using namespace gapx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[2048] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(GAPX_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                      \
    do {                            \
        int s_ = (x);               \
        if (s_ != GAPX_OK) return s_; \
    } while (0)

// ---------------------------------------------------------------------------
// minimal XML reader (elements, attributes, text, CDATA, comments)
// ---------------------------------------------------------------------------
struct XNode {
    std::string tag, text;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XNode>> kids;
    const char *attr(const char *k) const {
        for (auto &a : attrs)
            if (a.first == k) return a.second.c_str();
        return nullptr;
    }
    const XNode *child(const char *t) const {
        for (auto &c : kids)
            if (c->tag == t) return c.get();
        return nullptr;
    }
};

static std::string xml_unescape(const std::string &s) {
    std::string o;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '&') {
            size_t e = s.find(';', i);
            if (e != std::string::npos) {
                std::string ent = s.substr(i + 1, e - i - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                else o += s.substr(i, e - i + 1);
                i = e;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

static bool xml_parse(const char *p, const char *end, XNode &root, std::string &err) {
    std::vector<XNode *> stack{&root};
    while (p < end) {
        if (*p != '<') {
            const char *q = p;
            while (q < end && *q != '<') q++;
            stack.back()->text += xml_unescape(std::string(p, q));
            p = q;
            continue;
        }
        if (end - p >= 4 && !strncmp(p, "<!--", 4)) {
            const char *q = strstr(p, "-->");
            if (!q) { err = "unterminated comment"; return false; }
            p = q + 3;
            continue;
        }
        if (end - p >= 9 && !strncmp(p, "<![CDATA[", 9)) {
            const char *q = strstr(p, "]]>");
            if (!q) { err = "unterminated CDATA"; return false; }
            stack.back()->text += std::string(p + 9, q);
            p = q + 3;
            continue;
        }
        if (p[1] == '?' || p[1] == '!') {
            const char *q = (const char *)memchr(p, '>', end - p);
            if (!q) { err = "unterminated declaration"; return false; }
            p = q + 1;
            continue;
        }
        if (p[1] == '/') {
            const char *q = (const char *)memchr(p, '>', end - p);
            if (!q) { err = "unterminated end tag"; return false; }
            std::string name(p + 2, q);
            while (!name.empty() && isspace((unsigned char)name.back())) name.pop_back();
            if (stack.size() < 2 || stack.back()->tag != name) { err = "mismatched end tag </" + name + ">"; return false; }
            stack.pop_back();
            p = q + 1;
            continue;
        }
        const char *q = p + 1;
        while (q < end && !isspace((unsigned char)*q) && *q != '>' && *q != '/') q++;
        auto node = std::make_unique<XNode>();
        node->tag.assign(p + 1, q);
        bool selfclose = false;
        while (q < end) {
            while (q < end && isspace((unsigned char)*q)) q++;
            if (q >= end) break;
            if (*q == '/') { selfclose = true; q++; continue; }
            if (*q == '>') { q++; break; }
            const char *k0 = q;
            while (q < end && *q != '=' && !isspace((unsigned char)*q) && *q != '>') q++;
            std::string key(k0, q);
            while (q < end && isspace((unsigned char)*q)) q++;
            if (q < end && *q == '=') {
                q++;
                while (q < end && isspace((unsigned char)*q)) q++;
                char quote = *q;
                if (quote != '"' && quote != '\'') { err = "unquoted attribute " + key; return false; }
                const char *v0 = ++q;
                while (q < end && *q != quote) q++;
                node->attrs.emplace_back(key, xml_unescape(std::string(v0, q)));
                q++;
            }
        }
        XNode *raw = node.get();
        stack.back()->kids.push_back(std::move(node));
        if (!selfclose) stack.push_back(raw);
        p = q;
    }
    if (stack.size() != 1) { err = "unclosed element <" + stack.back()->tag + ">"; return false; }
    return true;
}

static void find_all(const XNode *n, const char *tag, std::vector<const XNode *> &out) {
    if (n->tag == tag) out.push_back(n);
    for (auto &c : n->kids) find_all(c.get(), tag, out);
}

// QUIP args-string parsing: key=value, key={a b c}, key="..." (ParamReader.f95:393-518)
static std::vector<std::pair<std::string, std::string>> parse_args(const std::string &s) {
    std::vector<std::pair<std::string, std::string>> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && isspace((unsigned char)s[i])) i++;
        if (i >= s.size()) break;
        size_t k0 = i;
        while (i < s.size() && s[i] != '=' && !isspace((unsigned char)s[i])) i++;
        std::string key = s.substr(k0, i - k0), val = "T";
        size_t j = i;
        while (j < s.size() && isspace((unsigned char)s[j])) j++;
        if (j < s.size() && s[j] == '=') {
            j++;
            while (j < s.size() && isspace((unsigned char)s[j])) j++;
            if (j < s.size() && (s[j] == '{' || s[j] == '"')) {
                char close = s[j] == '{' ? '}' : '"';
                size_t e = s.find(close, j + 1);
                if (e == std::string::npos) e = s.size();
                val = s.substr(j + 1, e - j - 1);
                i = e + 1;
            } else {
                size_t e = j;
                while (e < s.size() && !isspace((unsigned char)s[e])) e++;
                val = s.substr(j, e - j);
                i = e;
            }
        }
        out.emplace_back(key, val);
    }
    return out;
}
static const std::string *arg_get(const std::vector<std::pair<std::string, std::string>> &a, const char *k) {
    for (auto &p : a)
        if (p.first == k) return &p.second;
    return nullptr;
}
static std::vector<int> int_list(const std::string &s) {
    std::vector<int> v;
    const char *p = s.c_str();
    char *e;
    while (*p) {
        long x = strtol(p, &e, 10);
        if (e == p) { p++; continue; }
        v.push_back((int)x);
        p = e;
    }
    return v;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GAPX_OK;
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes + bytes / 4, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return fail(GAPX_E_DEVICE, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
        cap = want;
        return GAPX_OK;
    }
    template <class T> T *as() const { return (T *)p; }
    ~DBuf() { if (p) hipFree(p); }
};

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
struct Coord {
    int kind = 0;          // 1 soap, 2 distance_2b
    std::string label, desc;
    // soap
    double rc = 0, ctw = 0.5, sigma = 0.5, cw = 1.0, bee = 10.0;
    int n = 0, L = 0, S = 1, dim = 0;
    std::vector<int> species, centreZ;
    double u00 = 0;
    std::vector<double> rb, T;
    std::vector<int> qmap;
    DBuf d_rb, d_T, d_qmap;
    // d2b
    int Z1 = 0, Z2 = 0;
    // gp
    int cov = 2;
    double delta = 1, zeta = 1, signal_mean = 0;
    int M = 0;
    std::vector<double> Xs, alpha, scut, theta;
    DBuf d_Xs, d_w, d_theta;
    // mixed-precision copies of Xs (split hi/lo), built on demand
    int split_prec = -1;
    int dpad = 0, Mpad = 0;
    double wscale = 1.0;
    DBuf d_Xs_hi, d_Xs_lo, d_XsT_hi, d_XsT_lo;
};

struct Workspace {
    DBuf pos, Z, cell_of, map_shift, cell_cnt, cell_start, cell_fill, cell_atoms, cell_unsorted, scan_part,
        scan_part2;
    DBuf n1_cnt, n1_off, n1_j, n1_shift, n1_r, n1_rev, n2_cnt, n2_off, n2_fill, n2_i, n2_k, n2_is, n2_ks;
    DBuf nb_off, nb_j, nb_rev, nb_r, nb_d, pairF, pairE, esite, centres;
    DBuf X, norm, cbuf, W, Wf, e_part, G, Gparts, split_ws, split_ws2;
    DBuf f, le, lv, e, vir, block_part, minmax;
    std::vector<int> h_centres;
};

struct gapx_model {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string label;
    double e0[119] = {0};
    double E_scale = 1.0;
    double cutoff = 0;
    int precision = GAPX_PREC_FP64;
    std::vector<std::unique_ptr<Coord>> coords;
    DBuf d_e0;
    Workspace ws;
    bool timing = false;
    hipEvent_t ev[8] = {};
    double last_ms[6] = {0};
    int n1_total_last = 0;
};

struct gapx_nl {
    int N = 0, cells[3] = {0, 0, 0};
    std::vector<int> n1_off, n1_j, n1_shift, n2_off, n2_j, n2_idx, map_shift;
    std::vector<double> n1_r;
};

// SOAP radial basis: grid r_a = a*rb/n with rb = cutoff + sigma*sqrt(2 E ln 10),
// Gaussian-interpolation matrix C, overlap S = L L^T, transform T = C^-1 L.
static void chol_lower(int n, const std::vector<double> &S, std::vector<double> &L) {
    L.assign(n * n, 0.0);
    for (int j = 0; j < n; j++) {
        double s = S[j * n + j];
        for (int k = 0; k < j; k++) s -= L[j * n + k] * L[j * n + k];
        L[j * n + j] = std::sqrt(s);
        for (int i = j + 1; i < n; i++) {
            double t = S[i * n + j];
            for (int k = 0; k < j; k++) t -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = t / L[j * n + j];
        }
    }
}
static int soap_basis(Coord &c) {
    const int n = c.n;
    const double alpha = 0.5 / (c.sigma * c.sigma);
    const double cb = c.rc + c.sigma * std::sqrt(2.0 * c.bee * std::log(10.0));
    c.rb.resize(n);
    for (int a = 0; a < n; a++) c.rb[a] = (double)a * cb / (double)n;
    std::vector<double> C(n * n), S(n * n), L, Lc;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double ri = c.rb[i], rj = c.rb[j];
            C[i * n + j] = std::exp(-alpha * (ri - rj) * (ri - rj));
            S[i * n + j] = std::exp(-alpha * (ri * ri + rj * rj)) * std::sqrt(2.0) * std::pow(alpha, 1.5) * (ri + rj) +
                           alpha * std::exp(-0.5 * alpha * (ri - rj) * (ri - rj)) * std::sqrt(M_PI) *
                               (1.0 + alpha * (ri + rj) * (ri + rj)) * (1.0 + std::erf(std::sqrt(alpha / 2.0) * (ri + rj)));
        }
    chol_lower(n, S, L);
    chol_lower(n, C, Lc);
    for (int i = 0; i < n; i++)
        if (!(L[i * n + i] > 0) || !(Lc[i * n + i] > 0))
            return fail(GAPX_E_MODEL, "soap: radial basis overlap not positive definite (n_max=%d)", n);
    c.T.assign(n * n, 0.0);
    std::vector<double> y(n);
    for (int b = 0; b < n; b++) {
        for (int i = 0; i < n; i++) {
            double t = L[i * n + b];
            for (int k = 0; k < i; k++) t -= Lc[i * n + k] * y[k];
            y[i] = t / Lc[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < n; k++) t -= Lc[k * n + i] * c.T[k * n + b];
            c.T[i * n + b] = t / Lc[i * n + i];
        }
    }
    c.u00 = L[0];
    const int K = c.S * n;
    c.qmap.clear();
    for (int ia = 0; ia < K; ia++)
        for (int jb = 0; jb <= ia; jb++)
            for (int l = 0; l <= c.L; l++) c.qmap.push_back((ia << 16) | (jb << 8) | l);
    return GAPX_OK;
}

static int read_sparse_file(const std::string &path, size_t count, std::vector<double> &out) {
    FILE *fp = fopen(path.c_str(), "r");
    if (!fp) return fail(GAPX_E_IO, "cannot open sparseX file %s", path.c_str());
    out.resize(count);
    for (size_t i = 0; i < count; i++) {
        if (fscanf(fp, "%lf", &out[i]) != 1) {
            fclose(fp);
            return fail(GAPX_E_MODEL, "sparseX file %s: expected %zu values, read %zu", path.c_str(), count, i);
        }
    }
    fclose(fp);
    return GAPX_OK;
}

static int parse_descriptor(Coord &c, const std::string &s) {
    c.desc = s;
    std::string kind;
    size_t sp = s.find_first_of(" \t\n");
    kind = s.substr(0, sp);
    auto a = parse_args(sp == std::string::npos ? "" : s.substr(sp));
    auto num = [&](const char *k, double dflt) { auto v = arg_get(a, k); return v ? atof(v->c_str()) : dflt; };
    if (kind == "soap") {
        c.kind = 1;
        if (!arg_get(a, "cutoff") || !arg_get(a, "n_max") || !arg_get(a, "l_max"))
            return fail(GAPX_E_MODEL, "soap descriptor needs cutoff, n_max and l_max: '%s'", s.c_str());
        c.rc = num("cutoff", 0);
        c.ctw = num("cutoff_transition_width", 0.5);
        c.sigma = arg_get(a, "atom_sigma") ? num("atom_sigma", 0.5) : num("atom_gaussian_width", 0.5);
        c.cw = num("central_weight", 1.0);
        c.bee = num("basis_error_exponent", 10.0);
        c.n = (int)num("n_max", 0);
        c.L = (int)num("l_max", 0);
        c.S = (int)num("n_species", 1);
        int nZ = (int)num("n_Z", 1);
        c.species = arg_get(a, "species_Z") ? int_list(*arg_get(a, "species_Z")) : std::vector<int>(c.S, 0);
        c.centreZ = arg_get(a, "Z") ? int_list(*arg_get(a, "Z")) : std::vector<int>(nZ, 0);
        if ((int)c.species.size() != c.S || (int)c.centreZ.size() != nZ)
            return fail(GAPX_E_MODEL, "soap: n_species/species_Z or n_Z/Z mismatch in '%s'", s.c_str());
        if (c.S > kMaxSpecies || c.S < 1) return fail(GAPX_E_UNSUPPORTED, "soap: n_species=%d unsupported", c.S);
        if (!soap_supported(c.n, c.L))
            return fail(GAPX_E_UNSUPPORTED, "soap: n_max=%d l_max=%d not built (n_max<=16, l_max in 2..8,10,12)", c.n, c.L);
        if (c.S * c.n > 255 || c.L > 255) return fail(GAPX_E_UNSUPPORTED, "soap: too many radial channels");
        const int K = c.S * c.n;
        c.dim = K * (K + 1) / 2 * (c.L + 1) + 1;
        return soap_basis(c);
    }
    if (kind == "distance_2b") {
        c.kind = 2;
        c.rc = num("cutoff", 0);
        c.ctw = num("cutoff_transition_width", 0.5);
        c.Z1 = (int)num("Z1", 0);
        c.Z2 = (int)num("Z2", 0);
        c.dim = 1;
        return GAPX_OK;
    }
    return fail(GAPX_E_UNSUPPORTED, "descriptor type '%s' is not supported", kind.c_str());
}

static int upload(DBuf &b, const void *h, size_t bytes, hipStream_t s = nullptr) {
    CHK(b.ensure(bytes));
    if (s) HIPCHK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
    else HIPCHK(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
    return GAPX_OK;
}

extern "C" int gapx_api_version(void) { return GAPX_API_VERSION; }

extern "C" void gapx_last_error(char *buf, int len) {
    if (!buf || len <= 0) return;
    strncpy(buf, g_err, len - 1);
    buf[len - 1] = 0;
}

extern "C" int gapx_device_count(int *count) {
    if (!count) return fail(GAPX_E_ARG, "count is NULL");
    HIPCHK(hipGetDeviceCount(count));
    return GAPX_OK;
}

static int model_load(const char *xml_text, int64_t xml_len, const char *label, const char *xml_dir, int device,
                      gapx_model *m) {
    if (!xml_text) return fail(GAPX_E_ARG, "xml_text is NULL");
    if (xml_len < 0) xml_len = (int64_t)strlen(xml_text);
    XNode root;
    std::string err;
    if (!xml_parse(xml_text, xml_text + xml_len, root, err)) return fail(GAPX_E_MODEL, "XML: %s", err.c_str());
    std::vector<const XNode *> params;
    find_all(&root, "GAP_params", params);
    const XNode *gp = nullptr;
    for (auto p : params) {
        const char *l = p->attr("label");
        if (!label || !*label || (l && !strcmp(l, label))) { gp = p; break; }
    }
    if (!gp) return fail(GAPX_E_MODEL, "GAP_params with label '%s' not found", label ? label : "");
    m->label = gp->attr("label") ? gp->attr("label") : "";
    const char *ver = gp->attr("gap_version");
    if (ver && atoll(ver) > 1527075646LL)
        return fail(GAPX_E_MODEL, "gap_version %s is newer than this reader (1527075646)", ver);
    if (const XNode *data = gp->child("GAP_data"))
        for (auto &e : data->kids)
            if (e->tag == "e0" && e->attr("Z") && e->attr("value")) {
                int z = atoi(e->attr("Z"));
                if (z >= 0 && z < 119) m->e0[z] = atof(e->attr("value"));
            }
    const XNode *sparse = gp->child("gpSparse");
    if (!sparse) return fail(GAPX_E_MODEL, "no gpSparse in GAP_params");
    std::string dir = xml_dir && *xml_dir ? xml_dir : ".";
    for (auto &g : sparse->kids) {
        if (g->tag != "gpCoordinates") continue;
        auto c = std::make_unique<Coord>();
        c->label = g->attr("label") ? g->attr("label") : "";
        const XNode *dn = g->child("descriptor");
        if (!dn) return fail(GAPX_E_MODEL, "gpCoordinates %s: no descriptor", c->label.c_str());
        std::string ds = dn->text;
        size_t b0 = ds.find_first_not_of(" \t\r\n"), b1 = ds.find_last_not_of(" \t\r\n");
        ds = b0 == std::string::npos ? "" : ds.substr(b0, b1 - b0 + 1);
        CHK(parse_descriptor(*c, ds));
        const char *dims = g->attr("dimensions"), *nsx = g->attr("n_sparseX"), *fn = g->attr("sparseX_filename");
        const char *sv = g->attr("signal_variance"), *ct = g->attr("covariance_type");
        if (!dims || !nsx || !fn || !sv || !ct) return fail(GAPX_E_MODEL, "gpCoordinates %s: missing attribute", c->label.c_str());
        if (atoi(dims) != c->dim)
            return fail(GAPX_E_MODEL, "gpCoordinates %s: dimensions=%s but the descriptor has %d", c->label.c_str(), dims, c->dim);
        c->M = atoi(nsx);
        c->delta = atof(sv);
        c->cov = atoi(ct);
        c->signal_mean = g->attr("signal_mean") ? atof(g->attr("signal_mean")) : 0.0;
        c->zeta = g->attr("zeta") ? atof(g->attr("zeta")) : 1.0;
        if (c->cov != 1 && c->cov != 2) return fail(GAPX_E_UNSUPPORTED, "covariance_type %d not supported", c->cov);
        if (c->kind == 1 && c->cov != 2) return fail(GAPX_E_UNSUPPORTED, "soap with covariance_type %d", c->cov);
        if (c->kind == 2 && c->cov != 1) return fail(GAPX_E_UNSUPPORTED, "distance_2b with covariance_type %d", c->cov);
        c->alpha.assign(c->M, 0.0);
        c->scut.assign(c->M, 1.0);
        int seen = 0;
        for (auto &sx : g->kids) {
            if (sx->tag != "sparseX") continue;
            int i = sx->attr("i") ? atoi(sx->attr("i")) : 0;
            if (i < 1 || i > c->M) return fail(GAPX_E_MODEL, "sparseX index %d out of range", i);
            c->alpha[i - 1] = sx->attr("alpha") ? atof(sx->attr("alpha")) : 0.0;
            c->scut[i - 1] = sx->attr("sparseCutoff") ? atof(sx->attr("sparseCutoff")) : 1.0;
            seen++;
        }
        if (seen != c->M) return fail(GAPX_E_MODEL, "gpCoordinates %s: %d sparseX entries, n_sparseX=%d", c->label.c_str(), seen, c->M);
        if (const XNode *th = g->child("theta")) {
            const char *p = th->text.c_str();
            char *e;
            while (*p) {
                double v = strtod(p, &e);
                if (e == p) { p++; continue; }
                c->theta.push_back(v);
                p = e;
            }
        }
        if (c->kind == 2 && c->theta.size() != 1) return fail(GAPX_E_MODEL, "distance_2b needs one theta");
        CHK(read_sparse_file(dir + "/" + fn, (size_t)c->M * c->dim, c->Xs));
        std::vector<double> w(c->M);
        for (int s = 0; s < c->M; s++) w[s] = c->alpha[s] * c->delta * c->delta * c->scut[s];
        CHK(upload(c->d_Xs, c->Xs.data(), c->Xs.size() * 8));
        CHK(upload(c->d_w, w.data(), w.size() * 8));
        if (c->kind == 1) {
            CHK(upload(c->d_rb, c->rb.data(), c->rb.size() * 8));
            CHK(upload(c->d_T, c->T.data(), c->T.size() * 8));
            CHK(upload(c->d_qmap, c->qmap.data(), c->qmap.size() * 4));
        } else {
            CHK(upload(c->d_theta, c->theta.data(), c->theta.size() * 8));
        }
        m->cutoff = std::max(m->cutoff, c->rc);
        m->coords.push_back(std::move(c));
    }
    if (m->coords.empty()) return fail(GAPX_E_MODEL, "no gpCoordinates in gpSparse");
    CHK(upload(m->d_e0, m->e0, sizeof(m->e0)));
    return GAPX_OK;
}

extern "C" int gapx_model_load(const char *xml_text, int64_t xml_len, const char *label, const char *xml_dir,
                               int device, gapx_model **out) {
    if (!out) return fail(GAPX_E_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(GAPX_E_ARG, "device %d out of range (%d visible)", device, ndev);
    HIPCHK(hipSetDevice(device));
    auto m = new gapx_model();
    m->device = device;
    int st = model_load(xml_text, xml_len, label, xml_dir, device, m);
    if (st == GAPX_OK) {
        hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
        if (e != hipSuccess) st = fail(GAPX_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        for (int k = 0; k < 8 && st == GAPX_OK; k++)
            if (hipEventCreate(&m->ev[k]) != hipSuccess) st = fail(GAPX_E_DEVICE, "hipEventCreate failed");
    }
    if (st != GAPX_OK) {
        gapx_model_free(m);
        return st;
    }
    *out = m;
    return GAPX_OK;
}

extern "C" int gapx_model_free(gapx_model *m) {
    if (!m) return GAPX_OK;
    hipSetDevice(m->device);
    if (m->stream) hipStreamSynchronize(m->stream);
    for (int k = 0; k < 8; k++)
        if (m->ev[k]) hipEventDestroy(m->ev[k]);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return GAPX_OK;
}

extern "C" int gapx_model_cutoff(const gapx_model *m, double *cutoff) {
    if (!m || !cutoff) return fail(GAPX_E_ARG, "NULL argument");
    *cutoff = m->cutoff;
    return GAPX_OK;
}
extern "C" int gapx_model_set_e_scale(gapx_model *m, double s) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    m->E_scale = s;
    return GAPX_OK;
}
extern "C" int gapx_model_set_precision(gapx_model *m, int p) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (p != GAPX_PREC_FP64 && p != GAPX_PREC_BF16X3 && p != GAPX_PREC_FP16X3)
        return fail(GAPX_E_ARG, "unknown precision %d", p);
    m->precision = p;
    return GAPX_OK;
}
extern "C" int gapx_model_info(const gapx_model *m, int *n_coord, int coord, int *dim, int *n_sparse) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (n_coord) *n_coord = (int)m->coords.size();
    if (coord >= 0 && coord < (int)m->coords.size()) {
        if (dim) *dim = m->coords[coord]->dim;
        if (n_sparse) *n_sparse = m->coords[coord]->M;
    }
    return GAPX_OK;
}
extern "C" int gapx_model_set_timing(gapx_model *m, int on) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    m->timing = on != 0;
    return GAPX_OK;
}
extern "C" int gapx_last_timings(const gapx_model *m, double *ms) {
    if (!m || !ms) return fail(GAPX_E_ARG, "NULL argument");
    for (int k = 0; k < 6; k++) ms[k] = m->last_ms[k];
    return GAPX_OK;
}

// ---------------------------------------------------------------------------
// neighbour-list geometry (host 3x3 maths in the reference's operation order)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
static void cross3(const double *x, const double *y, double *o) {
    o[0] = +(x[1] * y[2] - x[2] * y[1]);
    o[1] = -(x[0] * y[2] - x[2] * y[0]);
    o[2] = +(x[0] * y[1] - x[1] * y[0]);
}
static double triple3(const double *x, const double *y, const double *z) {
    return +x[0] * (y[1] * z[2] - y[2] * z[1]) - x[1] * (y[0] * z[2] - y[2] * z[0]) + x[2] * (y[0] * z[1] - y[1] * z[0]);
}
static double norm3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void inv3(const double *m, double *g) {   // matrix3x3_inverse, linearalgebra.f95:6721-6733
    double stp = triple3(m, m + 3, m + 6), c[3];
    for (int r = 0; r < 3; r++) {
        cross3(m + 3 * ((r + 1) % 3), m + 3 * ((r + 2) % 3), c);
        for (int k = 0; k < 3; k++) g[3 * k + r] = c[k] / stp;
    }
}
static void mult3(const double *m, const double *v, double *o) {
    for (int i = 0; i < 3; i++) o[i] = 0.0;
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) o[i] = o[i] + m[3 * j + i] * v[j];
}

// minmax: fractional min[3], max[3] of the atoms (needed only with a non-periodic axis)
static void prepare_geom(const double *lat, const int *pbc, double cutoff, const double *minmax, NLGeom &g) {
    double eff[9], gi[9];
    memcpy(eff, lat, sizeof(eff));
    memcpy(g.lat, lat, sizeof(eff));
    for (int a = 0; a < 3; a++) { g.offset[a] = 0.0; g.pbc[a] = pbc[a] ? 1 : 0; }
    if (minmax)
        for (int i = 0; i < 3; i++) {
            if (pbc[i]) continue;
            double mn = minmax[i], mx = minmax[3 + i];
            if (mn <= 0.5 || mx >= 0.5) {   // calc_lat_eff, Connection.f95:740-763
                double ext = (mx - mn) * 1.1, nv = norm3(eff + 3 * i);
                for (int k = 0; k < 3; k++) eff[3 * i + k] = eff[3 * i + k] / nv * ext;
                g.offset[i] = mn - 0.05;
            }
        }
    inv3(eff, g.inv_eff);
    double bc[3], vol = std::fabs(triple3(eff, eff + 3, eff + 6));
    for (int a = 0; a < 3; a++) {   // divide_cell
        cross3(eff + 3 * ((a + 1) % 3), eff + 3 * ((a + 2) % 3), bc);
        g.cells[a] = std::max(1, (int)(vol / (cutoff * norm3(bc))));
    }
    inv3(lat, gi);   // fit_box_in_cell on the true lattice
    double mxs[3] = {0, 0, 0}, lc[3], corner[3];
    for (int k = -1; k <= 1; k += 2)
        for (int j = -1; j <= 1; j += 2)
            for (int i = -1; i <= 1; i += 2) {
                corner[0] = i * cutoff; corner[1] = j * cutoff; corner[2] = k * cutoff;
                mult3(gi, corner, lc);
                for (int a = 0; a < 3; a++) mxs[a] = std::max(mxs[a], std::fabs(lc[a]));
            }
    for (int a = 0; a < 3; a++) g.img[a] = std::max(1, ((int)std::ceil(2.0 * mxs[a]) + 1) / 2);
    g.cutoff = cutoff;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// pipeline
// ---------------------------------------------------------------------------
static int build_nl(gapx_model *m, int N, const double *d_pos, const double *lat, const int *pbc, double cutoff,
                    int *P_out) {
    Workspace &w = m->ws;
    hipStream_t s = m->stream;
    if (!(cutoff > 0)) return fail(GAPX_E_GEOMETRY, "calc_connect: cutoff must be positive (got %g)", cutoff);
    double vol = std::fabs(triple3(lat, lat + 3, lat + 6));
    if (!(vol > 0)) return fail(GAPX_E_GEOMETRY, "lattice is singular");
    NLGeom g;
    bool any_np = !pbc[0] || !pbc[1] || !pbc[2];
    double mm[6];
    if (any_np && N > 0) {
        double gi[9];
        inv3(lat, gi);
        CHK(w.minmax.ensure(16 * 8));
        CHK(w.split_ws.ensure(9 * 8));
        HIPCHK(hipMemcpyAsync(w.split_ws.p, gi, 9 * 8, hipMemcpyHostToDevice, s));
        launch_minmax_frac(N, d_pos, w.split_ws.as<double>(), w.minmax.as<double>(), s);
        HIPCHK(hipMemcpyAsync(mm, w.minmax.p, 6 * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    prepare_geom(lat, pbc, cutoff, any_np ? mm : nullptr, g);
    long ncell = (long)g.cells[0] * g.cells[1] * g.cells[2];
    if (ncell > (1L << 28)) return fail(GAPX_E_GEOMETRY, "too many cells (%ld)", ncell);
    const int NC = (int)ncell;
    CHK(w.cell_of.ensure((size_t)N * 4 + 4));
    CHK(w.map_shift.ensure((size_t)N * 12 + 12));
    CHK(w.cell_cnt.ensure((size_t)NC * 4));
    CHK(w.cell_fill.ensure((size_t)NC * 4));
    CHK(w.cell_start.ensure((size_t)(NC + 1) * 4));
    CHK(w.cell_atoms.ensure((size_t)N * 4 + 4));
    CHK(w.scan_part.ensure((size_t)scan_partials_size(std::max(NC, N) + 1) * 4 + 64));
    CHK(w.scan_part2.ensure((size_t)scan_partials_size(N + 1) * 4 + 64));
    CHK(w.n1_cnt.ensure((size_t)N * 4 + 4));
    CHK(w.n1_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n2_cnt.ensure((size_t)N * 4 + 4));
    CHK(w.n2_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n2_fill.ensure((size_t)N * 4 + 4));
    CHK(w.nb_off.ensure((size_t)(N + 1) * 4));
    CHK(w.cell_unsorted.ensure((size_t)N * 4 + 4));
    // per-call counters start at zero
    launch_fill_i32(w.cell_cnt.as<int>(), NC, 0, s);
    launch_fill_i32(w.cell_fill.as<int>(), NC, 0, s);
    launch_fill_i32(w.n2_cnt.as<int>(), N, 0, s);
    launch_fill_i32(w.n2_fill.as<int>(), N, 0, s);
    launch_nl_bin(N, d_pos, g, w.cell_of.as<int>(), w.map_shift.as<int>(), s);
    launch_nl_cell_count(N, w.cell_of.as<int>(), w.cell_cnt.as<int>(), s);
    launch_scan_i32(w.cell_cnt.as<int>(), NC, w.cell_start.as<int>(), w.scan_part.as<int>(), s);
    launch_nl_cell_fill(N, w.cell_of.as<int>(), w.cell_start.as<int>(), w.cell_fill.as<int>(),
                        w.cell_unsorted.as<int>(), s);
    launch_nl_cell_rank(N, w.cell_of.as<int>(), w.cell_start.as<int>(), w.cell_cnt.as<int>(),
                        w.cell_unsorted.as<int>(), w.cell_atoms.as<int>(), s);
    launch_nl_pairs(false, N, d_pos, w.map_shift.as<int>(), w.cell_of.as<int>(), w.cell_start.as<int>(),
                    w.cell_cnt.as<int>(), w.cell_atoms.as<int>(), g, w.n1_cnt.as<int>(), nullptr, nullptr, nullptr,
                    nullptr, w.n2_cnt.as<int>(), nullptr, nullptr, nullptr, nullptr, s);
    launch_scan_i32(w.n1_cnt.as<int>(), N, w.n1_off.as<int>(), w.scan_part.as<int>(), s);
    launch_scan_i32(w.n2_cnt.as<int>(), N, w.n2_off.as<int>(), w.scan_part2.as<int>(), s);
    int n1_total = 0;
    HIPCHK(hipMemcpyAsync(&n1_total, w.n1_off.as<int>() + N, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (n1_total < 0) return fail(GAPX_E_GEOMETRY, "neighbour count overflow");
    const size_t T1 = (size_t)n1_total + 1;
    CHK(w.n1_j.ensure(T1 * 4));
    CHK(w.n1_shift.ensure(T1 * 12));
    CHK(w.n1_r.ensure(T1 * 8));
    CHK(w.n1_rev.ensure(T1 * 4));
    CHK(w.n2_i.ensure(T1 * 4));
    CHK(w.n2_k.ensure(T1 * 4));
    CHK(w.n2_is.ensure(T1 * 4));
    CHK(w.n2_ks.ensure(T1 * 4));
    const size_t P = 2 * (size_t)n1_total + 1;
    CHK(w.nb_j.ensure(P * 4));
    CHK(w.nb_rev.ensure(P * 4));
    CHK(w.nb_r.ensure(P * 8));
    CHK(w.nb_d.ensure(P * 24));
    launch_nl_pairs(true, N, d_pos, w.map_shift.as<int>(), w.cell_of.as<int>(), w.cell_start.as<int>(),
                    w.cell_cnt.as<int>(), w.cell_atoms.as<int>(), g, nullptr, w.n1_off.as<int>(), w.n1_j.as<int>(),
                    w.n1_shift.as<int>(), w.n1_r.as<double>(), nullptr, w.n2_off.as<int>(), w.n2_fill.as<int>(),
                    w.n2_i.as<int>(), w.n2_k.as<int>(), s);
    launch_nl_n2_sort(N, w.n2_off.as<int>(), w.n2_i.as<int>(), w.n2_k.as<int>(), w.cell_of.as<int>(),
                      w.n1_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(), w.n1_rev.as<int>(), s);
    launch_nl_full(N, d_pos, g, w.n1_off.as<int>(), w.n1_j.as<int>(), w.n1_shift.as<int>(), w.n1_r.as<double>(),
                   w.n2_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(), w.n1_rev.as<int>(), w.nb_off.as<int>(),
                   w.nb_j.as<int>(), w.nb_rev.as<int>(), w.nb_r.as<double>(), w.nb_d.as<double>(), s);
    m->n1_total_last = n1_total;
    *P_out = (int)(P - 1);
    return GAPX_OK;
}

static bool z_in(const std::vector<int> &v, int z) {
    for (int x : v)
        if (x == z || x == 0) return true;
    return false;
}

// Scratch outputs.  Growing them frees the old blocks, so callers that hand
// these buffers in as outputs size them first.
static int ensure_outputs(Workspace &w, int N) {
    CHK(w.f.ensure((size_t)N * 24 + 24));
    CHK(w.le.ensure((size_t)N * 8 + 8));
    CHK(w.lv.ensure((size_t)N * 72 + 72));
    CHK(w.e.ensure(16));
    CHK(w.vir.ensure(80));
    return GAPX_OK;
}

static SoapDev soap_dev(const Coord &c) {
    SoapDev sp;
    sp.rc = c.rc; sp.ctw = c.ctw; sp.alpha = 0.5 / (c.sigma * c.sigma); sp.cw = c.cw; sp.u00 = c.u00;
    sp.n = c.n; sp.S = c.S; sp.K = c.S * c.n; sp.L = c.L; sp.LM = (c.L + 1) * (c.L + 1); sp.dim = c.dim;
    for (int s = 0; s < kMaxSpecies; s++) sp.species_Z[s] = s < c.S ? c.species[s] : -1;
    sp.rb = c.d_rb.as<double>(); sp.T = c.d_T.as<double>(); sp.qmap = c.d_qmap.as<int>();
    return sp;
}

static int ensure_split(gapx_model *m, Coord &c, int prec);

static void tmark(gapx_model *m, int k) {
    if (m->timing) hipEventRecord(m->ev[k], m->stream);
}

// Device-resident evaluation of one structure.
static int eval_device(gapx_model *m, int N, const double *d_pos, const int *Z, const double *lat, const int *pbc,
                       double nl_cutoff, unsigned want, double *d_e, double *d_le, double *d_f, double *d_v,
                       double *d_lv) {
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    if (N > 0 && (!d_pos || !Z)) return fail(GAPX_E_ARG, "positions / Z are NULL");
    if (!lat || !pbc) return fail(GAPX_E_ARG, "lattice / pbc are NULL");
    for (int i = 0; i < N; i++)
        if (Z[i] < 0 || Z[i] > 118) return fail(GAPX_E_ARG, "Z[%d]=%d out of range", i, Z[i]);
    Workspace &w = m->ws;
    hipStream_t s = m->stream;
    HIPCHK(hipSetDevice(m->device));
    const double rc = nl_cutoff > 0 ? nl_cutoff : m->cutoff;
    if (rc < m->cutoff) return fail(GAPX_E_ARG, "nl_cutoff %g is below the model cutoff %g", rc, m->cutoff);
    const bool grad = (want & (GAPX_WANT_F | GAPX_WANT_V | GAPX_WANT_LOCAL_V)) != 0;
    CHK(w.Z.ensure((size_t)N * 4 + 4));
    if (N) HIPCHK(hipMemcpyAsync(w.Z.p, Z, (size_t)N * 4, hipMemcpyHostToDevice, s));
    tmark(m, 0);
    int P = 0;
    CHK(build_nl(m, N, d_pos, lat, pbc, rc, &P));
    tmark(m, 1);
    CHK(w.pairF.ensure((size_t)(P + 1) * 24));
    CHK(w.pairE.ensure((size_t)(P + 1) * 8));
    CHK(w.esite.ensure((size_t)(N + 1) * 8));
    launch_fill_f64(w.pairF.as<double>(), (size_t)P * 3, 0.0, s);
    launch_fill_f64(w.pairE.as<double>(), (size_t)P, 0.0, s);
    launch_fill_f64(w.esite.as<double>(), (size_t)N, 0.0, s);
    double tsum[6] = {0, 0, 0, 0, 0, 0};
    for (auto &cp : m->coords) {
        Coord &c = *cp;
        if (c.kind == 2) {
            launch_d2b(N, w.Z.as<int>(), c.Z1, c.Z2, c.rc, c.ctw, c.M, c.d_Xs.as<double>(), c.d_w.as<double>(),
                       c.d_theta.as<double>(), w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_r.as<double>(),
                       w.nb_d.as<double>(), w.pairF.as<double>(), w.pairE.as<double>(), s);
            continue;
        }
        w.h_centres.clear();
        for (int i = 0; i < N; i++)
            if (z_in(c.centreZ, Z[i])) w.h_centres.push_back(i);
        const int nc = (int)w.h_centres.size();
        if (nc == 0) continue;
        CHK(w.centres.ensure((size_t)nc * 4));
        HIPCHK(hipMemcpyAsync(w.centres.p, w.h_centres.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
        const int KLM = c.S * c.n * (c.L + 1) * (c.L + 1);
        CHK(w.X.ensure((size_t)nc * c.dim * 8));
        CHK(w.norm.ensure((size_t)nc * 8));
        CHK(w.cbuf.ensure((size_t)nc * KLM * 8));
        CHK(w.W.ensure((size_t)nc * c.M * 8));
        CHK(w.G.ensure((size_t)nc * c.dim * 8));
        CHK(w.e_part.ensure((size_t)nc * ((c.M + 31) / 32 + 1) * 8));
        SoapDev sp = soap_dev(c);
        hipEvent_t e0 = m->ev[2], e1 = m->ev[3], e2 = m->ev[4], e3 = m->ev[5], e4 = m->ev[6];
        if (m->timing) hipEventRecord(e0, s);
        launch_soap_fwd(sp, nc, w.centres.as<int>(), w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                        w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.norm.as<double>(),
                        w.cbuf.as<double>(), s);
        if (m->timing) hipEventRecord(e1, s);
        int ntile = 0;
        if (m->precision == GAPX_PREC_FP64) {
            launch_gemm1_f64(nc, c.M, c.dim, w.X.as<double>(), c.d_Xs.as<double>(), c.d_w.as<double>(), c.zeta,
                             w.W.as<double>(), w.e_part.as<double>(), &ntile, s);
        } else {
            CHK(ensure_split(m, c, m->precision));
            CHK(w.Wf.ensure((size_t)nc * c.Mpad * 4 + 64));
            CHK(w.split_ws.ensure((size_t)(nc + 63) / 64 * 64 * c.dpad * 4 + 64));
            launch_gemm1_split(m->precision, nc, c.M, c.dim, c.dpad, w.X.as<double>(), c.d_Xs_hi.p, c.d_Xs_lo.p,
                               c.d_w.as<double>(), c.zeta, c.wscale, w.Wf.as<float>(), w.e_part.as<double>(), &ntile,
                               w.split_ws.p, s);
        }
        launch_energy_rows(nc, ntile, w.centres.as<int>(), w.e_part.as<double>(), w.esite.as<double>(), s);
        if (m->timing) hipEventRecord(e2, s);
        if (grad) {
            if (m->precision == GAPX_PREC_FP64) {
                CHK(w.Gparts.ensure((size_t)gemm2_parts(nc, c.dim) * nc * c.dim * 8 + 8));
                launch_gemm2_f64(nc, c.M, c.dim, w.W.as<double>(), c.d_Xs.as<double>(), w.G.as<double>(),
                                 w.Gparts.as<double>(), s);
            }
            else
                launch_gemm2_split(m->precision, nc, c.M, c.dim, c.dpad, c.Mpad, w.Wf.as<float>(), c.d_XsT_hi.p,
                                   c.d_XsT_lo.p, c.wscale, w.G.as<double>(), s);
            if (m->timing) hipEventRecord(e3, s);
            launch_soap_bwd(sp, nc, w.centres.as<int>(), w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                            w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.G.as<double>(),
                            w.norm.as<double>(), w.cbuf.as<double>(), w.pairF.as<double>(), s);
            if (m->timing) hipEventRecord(e4, s);
        }
        if (m->timing) {
            HIPCHK(hipStreamSynchronize(s));
            float t;
            hipEventElapsedTime(&t, e0, e1); tsum[1] += t;
            hipEventElapsedTime(&t, e1, e2); tsum[2] += t;
            if (grad) {
                hipEventElapsedTime(&t, e2, e3); tsum[3] += t;
                hipEventElapsedTime(&t, e3, e4); tsum[4] += t;
            }
        }
    }
    tmark(m, 7);
    // outputs not requested go to scratch (already sized when they are the caller's outputs)
    CHK(ensure_outputs(w, N));
    int nblk = 0;
    CHK(w.block_part.ensure((size_t)((N + 3) / 4 + 1) * 80));
    launch_gather(N, w.Z.as<int>(), m->d_e0.as<double>(), m->E_scale, w.nb_off.as<int>(), w.nb_rev.as<int>(),
                  w.nb_d.as<double>(), w.pairF.as<double>(), w.pairE.as<double>(), w.esite.as<double>(),
                  (want & GAPX_WANT_F) && d_f ? d_f : w.f.as<double>(),
                  (want & GAPX_WANT_LOCAL_E) && d_le ? d_le : w.le.as<double>(),
                  (want & GAPX_WANT_LOCAL_V) && d_lv ? d_lv : w.lv.as<double>(), w.block_part.as<double>(), &nblk, s);
    launch_final_reduce(nblk, w.block_part.as<double>(), m->E_scale,
                        (want & GAPX_WANT_E) && d_e ? d_e : w.e.as<double>(),
                        (want & GAPX_WANT_V) && d_v ? d_v : w.vir.as<double>(), s);
    HIPCHK(hipGetLastError());
    if (m->timing) {
        tmark(m, 6);
        HIPCHK(hipStreamSynchronize(s));
        float t;
        hipEventElapsedTime(&t, m->ev[0], m->ev[1]); tsum[0] = t;
        hipEventElapsedTime(&t, m->ev[7], m->ev[6]); tsum[5] = t;
        for (int k = 0; k < 6; k++) m->last_ms[k] = tsum[k];
    }
    return GAPX_OK;
}

extern "C" int gapx_calc_device(gapx_model *m, int N, const double *d_pos, const int *Z, const double *lat,
                                const int *pbc, double nl_cutoff, unsigned want, double *d_e, double *d_le,
                                double *d_f, double *d_v, double *d_lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    return eval_device(m, N, d_pos, Z, lat, pbc, nl_cutoff, want, d_e, d_le, d_f, d_v, d_lv);
}

extern "C" int gapx_sync(gapx_model *m) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    HIPCHK(hipStreamSynchronize(m->stream));
    return GAPX_OK;
}

extern "C" int gapx_calc(gapx_model *m, int N, const double *pos, const int *Z, const double *lat, const int *pbc,
                         double nl_cutoff, unsigned want, double *e, double *le, double *f, double *v, double *lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    if ((want & GAPX_WANT_E) && !e) return fail(GAPX_E_ARG, "energy requested but e is NULL");
    if ((want & GAPX_WANT_F) && !f) return fail(GAPX_E_ARG, "forces requested but f is NULL");
    if ((want & GAPX_WANT_V) && !v) return fail(GAPX_E_ARG, "virial requested but virial is NULL");
    if ((want & GAPX_WANT_LOCAL_E) && !le) return fail(GAPX_E_ARG, "local_e requested but NULL");
    if ((want & GAPX_WANT_LOCAL_V) && !lv) return fail(GAPX_E_ARG, "local_virial requested but NULL");
    HIPCHK(hipSetDevice(m->device));
    Workspace &w = m->ws;
    CHK(w.pos.ensure((size_t)N * 24 + 24));
    CHK(ensure_outputs(w, N));   // before taking the output pointers below
    if (N > 0) HIPCHK(hipMemcpyAsync(w.pos.p, pos, (size_t)N * 24, hipMemcpyHostToDevice, m->stream));
    CHK(eval_device(m, N, w.pos.as<double>(), Z, lat, pbc, nl_cutoff, want | GAPX_WANT_E | GAPX_WANT_V, w.e.as<double>(),
                    w.le.as<double>(), w.f.as<double>(), w.vir.as<double>(), w.lv.as<double>()));
    hipStream_t s = m->stream;
    if (want & GAPX_WANT_E) HIPCHK(hipMemcpyAsync(e, w.e.p, 8, hipMemcpyDeviceToHost, s));
    if (want & GAPX_WANT_V) HIPCHK(hipMemcpyAsync(v, w.vir.p, 72, hipMemcpyDeviceToHost, s));
    if ((want & GAPX_WANT_F) && N) HIPCHK(hipMemcpyAsync(f, w.f.p, (size_t)N * 24, hipMemcpyDeviceToHost, s));
    if ((want & GAPX_WANT_LOCAL_E) && N) HIPCHK(hipMemcpyAsync(le, w.le.p, (size_t)N * 8, hipMemcpyDeviceToHost, s));
    if ((want & GAPX_WANT_LOCAL_V) && N) HIPCHK(hipMemcpyAsync(lv, w.lv.p, (size_t)N * 72, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return GAPX_OK;
}

extern "C" int gapx_calc_batch(gapx_model *m, int n_struct, const int *atom_off, const double *pos, const int *Z,
                               const double *lattices, const int *pbcs, double nl_cutoff, unsigned want,
                               double *energies, double *local_e, double *f, double *virials, double *local_virial) {
    if (!m || !atom_off || n_struct < 0) return fail(GAPX_E_ARG, "bad batch arguments");
    for (int s = 0; s < n_struct; s++) {
        const int a0 = atom_off[s], n = atom_off[s + 1] - a0;
        if (n < 0) return fail(GAPX_E_ARG, "atom_off not ascending at %d", s);
        CHK(gapx_calc(m, n, pos + 3 * (size_t)a0, Z + a0, lattices + 9 * (size_t)s, pbcs + 3 * (size_t)s, nl_cutoff,
                      want, energies ? energies + s : nullptr, local_e ? local_e + a0 : nullptr,
                      f ? f + 3 * (size_t)a0 : nullptr, virials ? virials + 9 * (size_t)s : nullptr,
                      local_virial ? local_virial + 9 * (size_t)a0 : nullptr));
    }
    return GAPX_OK;
}

// ---------------------------------------------------------------------------
// neighbour list only
// ---------------------------------------------------------------------------
extern "C" int gapx_connect(int device, int N, const double *pos, const double *lat, const int *pbc, double cutoff,
                            gapx_nl **out) {
    if (!out || (N > 0 && !pos) || !lat || !pbc || N < 0) return fail(GAPX_E_ARG, "bad arguments");
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    gapx_model tmp;
    tmp.device = device;
    HIPCHK(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
    int st = tmp.ws.pos.ensure((size_t)N * 24 + 24);
    if (st == GAPX_OK && N > 0) {
        if (hipMemcpy(tmp.ws.pos.p, pos, (size_t)N * 24, hipMemcpyHostToDevice) != hipSuccess)
            st = fail(GAPX_E_DEVICE, "copy of positions failed");
    }
    int P = 0;
    if (st == GAPX_OK) st = build_nl(&tmp, N, tmp.ws.pos.as<double>(), lat, pbc, cutoff, &P);
    if (st == GAPX_OK) {
        auto nl = new gapx_nl();
        nl->N = N;
        const int T1 = tmp.n1_total_last;
        nl->n1_off.resize(N + 1); nl->n2_off.resize(N + 1); nl->map_shift.resize(3 * (size_t)N);
        nl->n1_j.resize(T1); nl->n1_shift.resize(3 * (size_t)T1); nl->n1_r.resize(T1);
        hipStreamSynchronize(tmp.stream);
        hipMemcpy(nl->n1_off.data(), tmp.ws.n1_off.p, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost);
        hipMemcpy(nl->n2_off.data(), tmp.ws.n2_off.p, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost);
        if (N) hipMemcpy(nl->map_shift.data(), tmp.ws.map_shift.p, (size_t)N * 12, hipMemcpyDeviceToHost);
        if (T1) {
            hipMemcpy(nl->n1_j.data(), tmp.ws.n1_j.p, (size_t)T1 * 4, hipMemcpyDeviceToHost);
            hipMemcpy(nl->n1_shift.data(), tmp.ws.n1_shift.p, (size_t)T1 * 12, hipMemcpyDeviceToHost);
            hipMemcpy(nl->n1_r.data(), tmp.ws.n1_r.p, (size_t)T1 * 8, hipMemcpyDeviceToHost);
        }
        const int T2 = nl->n2_off[N];
        nl->n2_j.resize(T2); nl->n2_idx.resize(T2);
        if (T2) {
            hipMemcpy(nl->n2_j.data(), tmp.ws.n2_is.p, (size_t)T2 * 4, hipMemcpyDeviceToHost);
            hipMemcpy(nl->n2_idx.data(), tmp.ws.n2_ks.p, (size_t)T2 * 4, hipMemcpyDeviceToHost);
        }
        NLGeom g;
        double mm[6] = {0, 0, 0, 0, 0, 0};
        prepare_geom(lat, pbc, cutoff, nullptr, g);
        for (int a = 0; a < 3; a++) nl->cells[a] = g.cells[a];
        (void)mm;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { st = fail(GAPX_E_DEVICE, "neighbour list: %s", hipGetErrorString(e)); delete nl; }
        else *out = nl;
    }
    hipStreamDestroy(tmp.stream);
    tmp.stream = nullptr;
    return st;
}
extern "C" int gapx_nl_counts(const gapx_nl *nl, int *n1_total, int *n2_total, int *cells) {
    if (!nl) return fail(GAPX_E_ARG, "NULL nl");
    if (n1_total) *n1_total = nl->n1_off[nl->N];
    if (n2_total) *n2_total = nl->n2_off[nl->N];
    if (cells) for (int a = 0; a < 3; a++) cells[a] = nl->cells[a];
    return GAPX_OK;
}
extern "C" int gapx_nl_export(const gapx_nl *nl, int *n1_off, int *n1_j, int *n1_shift, double *n1_r, int *n2_off,
                              int *n2_j, int *n2_idx, int *map_shift) {
    if (!nl) return fail(GAPX_E_ARG, "NULL nl");
    auto cp = [](void *dst, const void *src, size_t b) { if (dst && b) memcpy(dst, src, b); };
    cp(n1_off, nl->n1_off.data(), nl->n1_off.size() * 4);
    cp(n1_j, nl->n1_j.data(), nl->n1_j.size() * 4);
    cp(n1_shift, nl->n1_shift.data(), nl->n1_shift.size() * 4);
    cp(n1_r, nl->n1_r.data(), nl->n1_r.size() * 8);
    cp(n2_off, nl->n2_off.data(), nl->n2_off.size() * 4);
    cp(n2_j, nl->n2_j.data(), nl->n2_j.size() * 4);
    cp(n2_idx, nl->n2_idx.data(), nl->n2_idx.size() * 4);
    cp(map_shift, nl->map_shift.data(), nl->map_shift.size() * 4);
    return GAPX_OK;
}
extern "C" int gapx_nl_free(gapx_nl *nl) {
    delete nl;
    return GAPX_OK;
}

extern "C" int gapx_descriptors(gapx_model *m, int coord, int N, const double *pos, const int *Z, const double *lat,
                                const int *pbc, double nl_cutoff, double *X, int *desc_atom, int *n_desc) {
    if (!m || coord < 0 || coord >= (int)m->coords.size() || !X || !desc_atom || !n_desc || N < 0)
        return fail(GAPX_E_ARG, "bad arguments");
    Coord &c = *m->coords[coord];
    if (c.kind != 1) return fail(GAPX_E_UNSUPPORTED, "gapx_descriptors: coordinate %d is not soap", coord);
    HIPCHK(hipSetDevice(m->device));
    Workspace &w = m->ws;
    hipStream_t s = m->stream;
    CHK(w.pos.ensure((size_t)N * 24 + 24));
    if (N) HIPCHK(hipMemcpyAsync(w.pos.p, pos, (size_t)N * 24, hipMemcpyHostToDevice, s));
    CHK(w.Z.ensure((size_t)N * 4 + 4));
    if (N) HIPCHK(hipMemcpyAsync(w.Z.p, Z, (size_t)N * 4, hipMemcpyHostToDevice, s));
    int P = 0;
    CHK(build_nl(m, N, w.pos.as<double>(), lat, pbc, nl_cutoff > 0 ? nl_cutoff : c.rc, &P));
    w.h_centres.clear();
    for (int i = 0; i < N; i++)
        if (z_in(c.centreZ, Z[i])) w.h_centres.push_back(i);
    const int nc = (int)w.h_centres.size();
    *n_desc = nc;
    if (nc == 0) return GAPX_OK;
    CHK(w.centres.ensure((size_t)nc * 4));
    HIPCHK(hipMemcpyAsync(w.centres.p, w.h_centres.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
    const int KLM = c.S * c.n * (c.L + 1) * (c.L + 1);
    CHK(w.X.ensure((size_t)nc * c.dim * 8));
    CHK(w.norm.ensure((size_t)nc * 8));
    CHK(w.cbuf.ensure((size_t)nc * KLM * 8));
    launch_soap_fwd(soap_dev(c), nc, w.centres.as<int>(), w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                    w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.norm.as<double>(),
                    w.cbuf.as<double>(), s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(X, w.X.p, (size_t)nc * c.dim * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(desc_atom, w.h_centres.data(), (size_t)nc * 4);
    return GAPX_OK;
}

// mixed-precision sparse-point copies are built lazily the first time a model
// runs in a split precision
static int ensure_split(gapx_model *m, Coord &c, int prec) {
    if (c.split_prec == prec) return GAPX_OK;
    c.dpad = (c.dim + 31) / 32 * 32;
    c.Mpad = (c.M + 31) / 32 * 32;
    const size_t rows64 = (size_t)(c.M + 63) / 64 * 64;
    CHK(c.d_Xs_hi.ensure(rows64 * c.dpad * 2));
    CHK(c.d_Xs_lo.ensure(rows64 * c.dpad * 2));
    double wmax = 0.0;
    for (int k = 0; k < c.M; k++) wmax = std::max(wmax, std::fabs(c.alpha[k] * c.delta * c.delta * c.scut[k]));
    wmax *= std::max(1.0, c.zeta);
    c.wscale = (prec == GAPX_PREC_FP16X3 && wmax > 0) ? std::ldexp(1.0, 8 - (int)std::ceil(std::log2(wmax))) : 1.0;
    CHK(c.d_XsT_hi.ensure((size_t)c.Mpad * c.dpad * 2));
    CHK(c.d_XsT_lo.ensure((size_t)c.Mpad * c.dpad * 2));
    launch_split_matrix(prec, c.M, c.dim, c.dpad, c.d_Xs.as<double>(), c.dim, false, c.d_Xs_hi.p, c.d_Xs_lo.p,
                        m->stream);
    launch_split_matrix(prec, c.M, c.dim, c.Mpad, c.d_Xs.as<double>(), c.dim, true, c.d_XsT_hi.p, c.d_XsT_lo.p,
                        m->stream);
    HIPCHK(hipGetLastError());
    c.split_prec = prec;
    return GAPX_OK;
}
