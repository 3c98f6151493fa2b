// Host side of libgapx.so: the C ABI of include/gapx.h, the GAP XML model
// reader, device-resident model state and the per-call pipeline
//   neighbour list -> SOAP forward -> GEMM1 (K, fused GP epilogue) ->
//   GEMM2 (G) -> SOAP adjoint -> distance_2b -> deterministic gather.
#include <thread>

#include "gapx_host.h"

using namespace gapx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[2048] = "";

int gapx::fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
const char *gapx::last_error_text() { return g_err; }

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
        // the buffer is length-delimited and need not be NUL-terminated: every search
        // stays inside [p, end)
        if (end - p >= 4 && !strncmp(p, "<!--", 4)) {
            static const char tail[] = "-->";
            const char *q = std::search(p + 4, end, tail, tail + 3);
            if (q == end) { err = "unterminated comment"; return false; }
            p = q + 3;
            continue;
        }
        if (end - p >= 9 && !strncmp(p, "<![CDATA[", 9)) {
            static const char tail[] = "]]>";
            const char *q = std::search(p + 9, end, tail, tail + 3);
            if (q == end) { err = "unterminated CDATA"; return false; }
            stack.back()->text += std::string(p + 9, q);
            p = q + 3;
            continue;
        }
        if (end - p < 2) { err = "truncated tag"; return false; }
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
        bool selfclose = false, closed = false;
        while (q < end) {
            while (q < end && isspace((unsigned char)*q)) q++;
            if (q >= end) break;
            if (*q == '/') { selfclose = true; q++; continue; }
            if (*q == '>') { q++; closed = true; break; }
            const char *k0 = q;
            while (q < end && *q != '=' && !isspace((unsigned char)*q) && *q != '>') q++;
            std::string key(k0, q);
            while (q < end && isspace((unsigned char)*q)) q++;
            if (q < end && *q == '=') {
                q++;
                while (q < end && isspace((unsigned char)*q)) q++;
                if (q >= end) { err = "truncated attribute " + key; return false; }
                char quote = *q;
                if (quote != '"' && quote != '\'') { err = "unquoted attribute " + key; return false; }
                const char *v0 = ++q;
                while (q < end && *q != quote) q++;
                if (q >= end) { err = "unterminated value of attribute " + key; return false; }
                node->attrs.emplace_back(key, xml_unescape(std::string(v0, q)));
                q++;
            }
        }
        if (!closed) { err = "unterminated tag <" + node->tag + ">"; return false; }
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

// Fortran list-directed logical: T, .true., t... are true
static bool fortran_true(const std::string &v) {
    size_t i = 0;
    while (i < v.size() && (isspace((unsigned char)v[i]) || v[i] == '.')) i++;
    return i < v.size() && (v[i] == 'T' || v[i] == 't');
}

// Descriptor options that change the descriptor and are not built: each is accepted
// only at its default (GAP's descriptor parser ignores keys it does not know, since
// gap_fit's own keys ride in the same string, so unknown keys are ignored here too).
struct OptDefault { const char *key; bool logical; double value; };
static const OptDefault kSoapDefaults[] = {
    {"normalise", true, 1}, {"normalize", true, 1}, {"average", true, 0}, {"diagonal_radial", true, 0},
    {"central_reference_all_species", true, 0}, {"covariance_sigma0", false, 0.0}, {"cutoff_dexp", false, 0.0},
    {"cutoff_scale", false, 1.0}, {"cutoff_rate", false, 1.0}, {"radial_enhancement", false, 0.0},
    {"nu_R", false, 2.0}, {"nu_S", false, 2.0}, {"Z_mix", true, 0}, {"R_mix", true, 0}, {"sym_mix", true, 0},
    {"compress_file", false, 0.0}, {"compress_mode", false, 0.0}};
static const OptDefault kD2bDefaults[] = {
    {"only_intra", true, 0}, {"only_inter", true, 0}, {"n_exponents", false, 1.0}, {"exponents", false, 1.0},
    {"resid_name", false, 0.0}};

static int check_defaults(const std::vector<std::pair<std::string, std::string>> &a, const OptDefault *tab, int n,
                          const std::string &desc) {
    for (int q = 0; q < n; q++) {
        const std::string *v = arg_get(a, tab[q].key);
        if (!v) continue;
        bool ok;
        if (tab[q].logical) ok = fortran_true(*v) == (tab[q].value != 0);
        else if (!strcmp(tab[q].key, "compress_file") || !strcmp(tab[q].key, "compress_mode") ||
                 !strcmp(tab[q].key, "resid_name"))
            ok = v->empty();
        else {
            char *e = nullptr;
            const double x = strtod(v->c_str(), &e);
            ok = e != v->c_str() && x == tab[q].value;
        }
        if (!ok)
            return fail(GAPX_E_UNSUPPORTED, "descriptor option %s=%s is not built (only its default): '%s'",
                        tab[q].key, v->c_str(), desc.c_str());
    }
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
        CHK(check_defaults(a, kSoapDefaults, (int)(sizeof(kSoapDefaults) / sizeof(kSoapDefaults[0])), s));
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
        CHK(check_defaults(a, kD2bDefaults, (int)(sizeof(kD2bDefaults) / sizeof(kD2bDefaults[0])), s));
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

static SoapDev soap_dev(const Coord &c);

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
    if (const XNode *data = gp->child("GAP_data")) {
        // a scalar e0 attribute sets every species, then <e0 Z value> entries override
        // (IPModel_GAP_read_params_xml, IPModel_GAP.f95:670-674, 685-701)
        if (const char *e0 = data->attr("e0"))
            for (int z = 0; z < 119; z++) m->e0[z] = atof(e0);
        if (const char *pca = data->attr("do_pca"))
            if (fortran_true(pca)) return fail(GAPX_E_UNSUPPORTED, "GAP_data do_pca=T is not built");
        for (auto &e : data->kids)
            if (e->tag == "e0") {
                if (!e->attr("Z") || !e->attr("value")) return fail(GAPX_E_MODEL, "GAP_data e0 needs Z and value");
                int z = atoi(e->attr("Z"));
                if (z < 0 || z >= 119) return fail(GAPX_E_MODEL, "GAP_data e0 Z=%d out of range", z);
                m->e0[z] = atof(e->attr("value"));
            }
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
        if (c->signal_mean != 0.0)
            return fail(GAPX_E_UNSUPPORTED, "gpCoordinates %s: signal_mean=%s is not built (only 0)", c->label.c_str(),
                        g->attr("signal_mean"));
        if (const char *np = g->attr("n_permutations"))
            if (atoi(np) > 1)
                return fail(GAPX_E_UNSUPPORTED, "gpCoordinates %s: n_permutations=%s is not built (only 1)",
                            c->label.c_str(), np);
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
        m->cutoff = std::max(m->cutoff, c->rc);
        m->coords.push_back(std::move(c));
    }
    if (m->coords.empty()) return fail(GAPX_E_MODEL, "no gpCoordinates in gpSparse");
    return GAPX_OK;
}

// Device half of the model load, on the current device: sparse points, coefficients and
// the SOAP tables; refuses descriptor shapes whose kernels need more LDS than a block has.
static int model_upload(gapx_model *m) {
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device));
    for (auto &cp : m->coords) {
        Coord *c = cp.get();
        std::vector<double> w(c->M);
        for (int s = 0; s < c->M; s++) w[s] = c->alpha[s] * c->delta * c->delta * c->scut[s];
        CHK(upload(c->d_Xs, c->Xs.data(), c->Xs.size() * 8));
        CHK(upload(c->d_w, w.data(), w.size() * 8));
        if (c->kind == 1) {
            if (gp_f64_fused_supported(c->dim)) {
                const int dp = (c->dim + 31) / 32 * 32, mp = (c->M + 31) / 32 * 32;
                std::vector<double> xp((size_t)mp * dp, 0.0);
                for (int r = 0; r < c->M; r++)
                    std::copy(c->Xs.begin() + (size_t)r * c->dim, c->Xs.begin() + (size_t)(r + 1) * c->dim,
                              xp.begin() + (size_t)r * dp);
                CHK(upload(c->d_XsP, xp.data(), xp.size() * 8));
            }
            CHK(upload(c->d_rb, c->rb.data(), c->rb.size() * 8));
            CHK(upload(c->d_T, c->T.data(), c->T.size() * 8));
            CHK(upload(c->d_qmap, c->qmap.data(), c->qmap.size() * 4));
            const size_t need = soap_lds_bytes(soap_dev(*c));
            if (need > (size_t)lds)
                return fail(GAPX_E_UNSUPPORTED,
                            "soap n_max=%d l_max=%d with %d species needs %zu bytes of LDS per block (%d available)",
                            c->n, c->L, c->S, need, lds);
        } else {
            CHK(upload(c->d_theta, c->theta.data(), c->theta.size() * 8));
        }
    }
    CHK(upload(m->d_e0, m->e0, sizeof(m->e0)));
    return GAPX_OK;
}

int gapx::new_engine(gapx_model *m, std::unique_ptr<Engine> &out) {
    std::unique_ptr<Engine> e(new Engine());
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return fail(GAPX_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(he));
    e->ws.centres.resize(m->coords.size());
    e->ws.var_inst.resize(m->coords.size());
    e->ws.cut2_inst.resize(m->coords.size());
    e->h_centres.resize(m->coords.size());
    out = std::move(e);
    return GAPX_OK;
}

extern "C" int gapx_model_load(const char *xml_text, int64_t xml_len, const char *label, const char *xml_dir,
                               int device, gapx_model **out) {
    if (!out) return fail(GAPX_E_ARG, "out is NULL");
    *out = nullptr;
    // the XML and sparseX files are read and checked before any device is touched
    std::unique_ptr<gapx_model> parsed(new gapx_model());
    parsed->device = device;
    CHK(model_load(xml_text, xml_len, label, xml_dir, device, parsed.get()));
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(GAPX_E_ARG, "device %d out of range (%d visible)", device, ndev);
    HIPCHK(hipSetDevice(device));
    auto m = parsed.release();
    int st = model_upload(m);
    if (st == GAPX_OK) {
        const hipError_t ce = soap_constants_init();
        st = ce != hipSuccess ? fail(GAPX_E_DEVICE, "uploading the SOAP constant tables: %s", hipGetErrorString(ce))
                              : new_engine(m, m->eng);
    }
    if (st == GAPX_OK) st = new_engine(m, m->eng_na);
    if (st == GAPX_OK) {
        Engine &na = *m->eng_na;
        for (HBuf *h : {&na.stage_in, &na.stage_geo, &na.stage_io, &na.ctl}) h->pageable = true;
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
    m->eng.reset();
    m->eng_na.reset();
    delete m;
    return GAPX_OK;
}

void gapx::reset_arena_views(Engine &eg) {
    Workspace &w = eg.ws;
    for (DBuf *b : {&w.pos, &w.Z, &w.sid, &w.atom_off, &w.geo, &w.cell_base, &w.ginv, &w.cell_of, &w.map_shift,
                    &w.cell_cnt, &w.cell_start, &w.cell_fill, &w.cell_atoms, &w.cell_unsorted, &w.scan_part, &w.cpos, &w.cms,
                    &w.scan_part2, &w.n1_cnt, &w.n1_off, &w.n1_j, &w.n1_shift, &w.n1_r, &w.n1_rev, &w.n2_cnt,
                    &w.n2_off, &w.n2_fill, &w.n2_i, &w.n2_k, &w.n2_is, &w.n2_ks, &w.nb_off, &w.nb_j, &w.nb_rev,
                    &w.nb_r, &w.nb_d, &w.pairF, &w.pairE, &w.esite, &w.X, &w.norm, &w.cbuf, &w.W, &w.Wf, &w.e_part,
                    &w.G, &w.Gparts, &w.split_ws, &w.f, &w.le, &w.lv, &w.e, &w.vir, &w.minmax, &w.last_pos,
                    &w.skin_flags, &w.cmask, &w.t1_j, &w.t1_shift, &w.t1_r, &w.t2_i, &w.t2_k, &w.nl_flags, &w.eatom,
                    &w.vK, &w.vk, &w.vV, &w.Gv, &w.pairFv, &w.pairVE, &w.varsite, &w.o_epc, &w.o_lgv, &w.o_gvg,
                    &w.spart})
        b->drop_view();
    for (auto *v : {&w.centres, &w.var_inst, &w.cut2_inst})
        for (DBuf &b : *v) b.drop_view();
    // uploads and neighbour lists refer to the old views
    eg.have_inputs = eg.have_geo = eg.nl_valid = false;
    eg.var_valid = false;
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
    if (p != GAPX_PREC_FP64) {   // the split copies of the sparse points are made here, at set-up
        HIPCHK(hipSetDevice(m->device));
        for (auto &cp : m->coords)
            if (cp->kind == 1) CHK(ensure_split(m, *cp, p, m->eng->stream));
        HIPCHK(hipStreamSynchronize(m->eng->stream));
    }
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
extern "C" int gapx_model_set_nl_reuse(gapx_model *m, int on) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    m->nl_reuse = on != 0;
    return GAPX_OK;
}
extern "C" int gapx_last_nl_rebuilt(const gapx_model *m, int *rebuilt) {
    if (!m || !rebuilt) return fail(GAPX_E_ARG, "NULL argument");
    *rebuilt = m->eng->last_rebuilt ? 1 : 0;
    return GAPX_OK;
}
extern "C" int gapx_model_set_graphs(gapx_model *m, int on) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    m->graphs = on != 0;
    return GAPX_OK;
}
extern "C" int gapx_model_set_autotune(gapx_model *m, int on) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    m->autotune = on != 0;
    return GAPX_OK;
}
extern "C" int gapx_model_set_fused_schedule(gapx_model *m, int parts) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (parts < -1 || parts > 8) return fail(GAPX_E_ARG, "fused schedule %d: -1 (stream-K), 0 or 1..8 parts", parts);
    m->fused_schedule = parts;
    return GAPX_OK;
}
extern "C" int gapx_model_tuning(gapx_model *m, int coord, int max_entries, int *bucket, int *gemm1_tile,
                                 int *gemm2_tile, int *gemm2_split, double *gemm1_ms, double *gemm2_ms,
                                 int *n_entries) {
    if (!m || !n_entries) return fail(GAPX_E_ARG, "NULL argument");
    if (coord < 0 || coord >= (int)m->coords.size()) return fail(GAPX_E_ARG, "coordinate %d out of range", coord);
    Coord &c = *m->coords[coord];
    std::lock_guard<std::mutex> lk(c.tune_mu);
    int k = 0;
    for (auto &e : c.tune) {
        if (k < max_entries) {
            if (bucket) bucket[k] = e.first;
            const bool fz = e.second.f64parts > 0;   // the fused kernel: tile -1, its parts and time
            if (gemm1_tile) gemm1_tile[k] = fz ? -1 : e.second.g1;
            if (gemm2_tile) gemm2_tile[k] = fz ? -1 : e.second.g2;
            if (gemm2_split) gemm2_split[k] = fz ? e.second.f64parts : e.second.parts;
            if (gemm1_ms) gemm1_ms[k] = fz ? e.second.f64_ms : e.second.g1_ms;
            if (gemm2_ms) gemm2_ms[k] = fz ? 0.0 : e.second.g2_ms;
        }
        k++;
    }
    *n_entries = k;
    return GAPX_OK;
}

extern "C" int gapx_model_set_timing(gapx_model *m, int on) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipStreamSynchronize(m->eng->stream));
    m->eng->timer.reset();
    m->eng->timer.on = on != 0;
    return GAPX_OK;
}
extern "C" int gapx_last_timings(gapx_model *m, double *ms) {
    if (!m || !ms) return fail(GAPX_E_ARG, "NULL argument");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipStreamSynchronize(m->eng->stream));
    m->eng->timer.resolve_all();
    for (int k = 0; k < kStages; k++) ms[k] = m->eng->timer.last[k];
    return GAPX_OK;
}
extern "C" int gapx_timing_totals(gapx_model *m, double *ms, int64_t *ncalls) {
    if (!m || !ms) return fail(GAPX_E_ARG, "NULL argument");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipStreamSynchronize(m->eng->stream));
    m->eng->timer.resolve_all();
    for (int k = 0; k < kStages; k++) ms[k] = m->eng->timer.total[k];
    if (ncalls) *ncalls = m->eng->timer.ncalls;
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
    memcpy(g.ginv, gi, sizeof(gi));
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
// fractional min/max per axis of atoms [a0, a1) from host positions, in the
// same operation order as the device kernel (g .mult. pos)
#pragma clang fp contract(off)
static void host_minmax(const double *pos, int a0, int a1, const double *gi, double *mm) {
    for (int a = 0; a < 3; a++) { mm[a] = 1e300; mm[3 + a] = -1e300; }
    for (int n = a0; n < a1; n++) {
        double lp[3];
        mult3(gi, pos + 3 * (size_t)n, lp);
        for (int a = 0; a < 3; a++) { mm[a] = std::min(mm[a], lp[a]); mm[3 + a] = std::max(mm[3 + a], lp[a]); }
    }
}
#pragma clang fp contract(on)

// Pinned staging is reused call after call; in a speculative call nothing waits for the
// stream, so staging written while an earlier upload from it may be queued waits first.
static int staging_ready(Engine &eg) {
    if (eg.staging_dirty) {
        HIPCHK(hipStreamSynchronize(eg.stream));
        eg.staging_dirty = false;
    }
    return GAPX_OK;
}

static bool geo_equal(const NLGeom &x, const NLGeom &y) { return memcmp(&x, &y, sizeof(NLGeom)) == 0; }

static bool z_in(const std::vector<int> &v, int z) {
    for (int x : v)
        if (x == z || x == 0) return true;
    return false;
}

// Upload Z, the atom -> structure map and the structure offsets when they differ
// from the previous call's; rebuild the per-coordinate centre lists with them.
// local (nullable): only atoms with local[i] != 0 are descriptor centres
// (IPModel_GAP's atom_mask_name); the mask is uploaded too and never cached.
static int stage_inputs(gapx_model *m, Engine &eg, int S, const int *atom_off, const int *Z,
                        const unsigned char *local = nullptr) {
    const int N = atom_off[S];
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    bool same = !local && eg.have_inputs && (int)eg.h_off.size() == S + 1 && (int)eg.h_Z.size() == N &&
                std::equal(atom_off, atom_off + S + 1, eg.h_off.begin()) && std::equal(Z, Z + N, eg.h_Z.begin());
    if (same) return GAPX_OK;
    eg.have_inputs = false;
    eg.nl_valid = false;   // a stored neighbour list belongs to other atoms
    eg.h_off.assign(atom_off, atom_off + S + 1);
    eg.h_Z.assign(Z, Z + N);
    eg.h_sid.resize(N);
    for (int st = 0; st < S; st++)
        for (int i = atom_off[st]; i < atom_off[st + 1]; i++) eg.h_sid[i] = st;
    size_t tot = (size_t)N * 8 + (size_t)(S + 1) * 4;
    for (size_t k = 0; k < m->coords.size(); k++) {
        auto &hc = eg.h_centres[k];
        hc.clear();
        const Coord &c = *m->coords[k];
        if (c.kind == 1)
            for (int i = 0; i < N; i++)
                if (z_in(c.centreZ, Z[i]) && (!local || local[i])) hc.push_back(i);
        tot += hc.size() * 4;
    }
    if (local) tot += (size_t)N + 8;
    CHK(w.Z.ensure((size_t)N * 4 + 4));
    CHK(w.sid.ensure((size_t)N * 4 + 4));
    CHK(w.atom_off.ensure((size_t)(S + 1) * 4));
    CHK(staging_ready(eg));   // an earlier upload from the staging may still be queued
    CHK(eg.stage_in.ensure(tot + 64));
    char *h = eg.stage_in.as<char>();
    eg.staging_dirty = true;
    size_t o = 0;
    auto put = [&](DBuf &dst, const void *src, size_t bytes) -> int {
        if (!bytes) return GAPX_OK;
        memcpy(h + o, src, bytes);
        HIPCHK(hipMemcpyAsync(dst.p, h + o, bytes, hipMemcpyHostToDevice, s));
        o += bytes;
        return GAPX_OK;
    };
    CHK(put(w.Z, Z, (size_t)N * 4));
    CHK(put(w.sid, eg.h_sid.data(), (size_t)N * 4));
    CHK(put(w.atom_off, atom_off, (size_t)(S + 1) * 4));
    for (size_t k = 0; k < m->coords.size(); k++) {
        auto &hc = eg.h_centres[k];
        CHK(w.centres[k].ensure(hc.size() * 4 + 4));
        CHK(put(w.centres[k], hc.data(), hc.size() * 4));
    }
    if (local) {
        CHK(w.cmask.ensure((size_t)N + 8));
        CHK(put(w.cmask, local, (size_t)N));
    }
    eg.have_inputs = !local;   // a masked call's centre lists are not reused
    return GAPX_OK;
}

// calc_connect for S structures at once (Connection.f95:1035-1310): per-structure
// geometry, one set of kernels over all atoms; pairs never cross structures.
//
// Two ways to size the tables:
//  - exact: the candidate search counts (or fills per-atom slots sized from the last
//    build's largest counts), the neighbour1 total is read back with one host
//    synchronisation, and the tables are sized to it;
//  - speculative (spec): every table is sized from capacities kept from earlier builds
//    (per-atom slots, the neighbour1 total), nothing is read back, and the whole build can
//    sit in a captured graph.  A device guard empties every list when a slot or the
//    total overflows, so the kernels after it stay inside their buffers; the flags are
//    read back when the caller synchronises and an overflowed call is redone exactly.
//    For structures with a non-periodic axis, lat_eff comes from a device kernel on this
//    call's extents; the cell counts are the host's (from the previous call's extents),
//    and a structure whose counts would change flags the call for redoing as well.
static int ensure_nl_tables(Workspace &w, int T1cap);

struct NLPlan {
    int N = 0, S = 0, NC = 0;
    bool spec = false, one_pass = false, any_np = false, dev_geo = false;
    int cap1 = 0, cap2 = 0, capT1 = 0;
};

// the lattice-derived fields of two geometries (not lat_eff, which the device may refresh)
static bool geo_equal_static(const NLGeom &x, const NLGeom &y) {
    return !memcmp(x.lat, y.lat, sizeof(x.lat)) && !memcmp(x.ginv, y.ginv, sizeof(x.ginv)) &&
           !memcmp(x.cells, y.cells, sizeof(x.cells)) && !memcmp(x.img, y.img, sizeof(x.img)) &&
           !memcmp(x.pbc, y.pbc, sizeof(x.pbc)) && x.cutoff == y.cutoff;
}

// Host geometry, uploads and buffers of a build.  Extents of non-periodic axes: from
// host positions when the caller has them, else from the device (exact: read back now;
// speculative: the last read extents for the cell counts, and this call's own on the
// device for lat_eff).
static int nl_prepare(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *d_pos, const double *h_pos,
                      const double *lats, const int *pbcs, double cutoff, bool spec, NLPlan &pl) {
    (void)m;
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int N = atom_off[S];
    if (!(cutoff > 0)) return fail(GAPX_E_GEOMETRY, "calc_connect: cutoff must be positive (got %g)", cutoff);
    pl = NLPlan();
    pl.N = N;
    pl.S = S;
    std::vector<double> mm;
    bool any_np = false;
    for (int st = 0; st < S; st++) {
        const double *lat = lats + 9 * (size_t)st;
        const double vol = std::fabs(triple3(lat, lat + 3, lat + 6));
        if (!(vol > 0)) return fail(GAPX_E_GEOMETRY, "lattice of structure %d is singular", st);
        const int *pbc = pbcs + 3 * st;
        if ((!pbc[0] || !pbc[1] || !pbc[2]) && atom_off[st + 1] > atom_off[st]) any_np = true;
    }
    pl.any_np = any_np;
    // a speculative build needs the slot and table capacities of an earlier build, and for
    // non-periodic axes read from the device the last read extents of as many structures
    if (spec && (eg.nl_cap1 <= 0 || eg.nl_cap2 <= 0 || eg.nl_capT1 <= 0)) spec = false;
    if (spec && any_np && !h_pos && eg.h_mm.size() != 6 * (size_t)S) spec = false;
    const bool dev_geo = spec && any_np && !h_pos;
    if (any_np) {
        mm.assign(6 * (size_t)S, 0.0);
        std::vector<double> gi(9 * (size_t)S);
        for (int st = 0; st < S; st++) inv3(lats + 9 * (size_t)st, gi.data() + 9 * st);
        if (h_pos) {
            for (int st = 0; st < S; st++) host_minmax(h_pos, atom_off[st], atom_off[st + 1], gi.data() + 9 * st,
                                                       mm.data() + 6 * st);
        } else {
            CHK(w.ginv.ensure(gi.size() * 8));
            CHK(w.minmax.ensure(mm.size() * 8));
            if (eg.h_ginv != gi) {
                CHK(staging_ready(eg));
                CHK(eg.stage_geo.ensure(gi.size() * 8 + mm.size() * 8));
                memcpy(eg.stage_geo.p, gi.data(), gi.size() * 8);
                HIPCHK(hipMemcpyAsync(w.ginv.p, eg.stage_geo.p, gi.size() * 8, hipMemcpyHostToDevice, s));
                eg.staging_dirty = true;
                eg.h_ginv = gi;
            }
            if (dev_geo) {
                mm = eg.h_mm;   // cell counts from the last read extents; lat_eff from this call's (device)
            } else {
                CHK(eg.ctl.ensure(mm.size() * 8 + 64));
                launch_minmax_frac(S, w.atom_off.as<int>(), d_pos, w.ginv.as<double>(), w.minmax.as<double>(), s);
                HIPCHK(hipMemcpyAsync(eg.ctl.p, w.minmax.p, mm.size() * 8, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                eg.staging_dirty = false;
                memcpy(mm.data(), eg.ctl.p, mm.size() * 8);
                eg.h_mm = mm;
            }
        }
    }
    std::vector<NLGeom> geo(S);
    std::vector<int> cbase(S + 1);
    long ncell = 0;
    for (int st = 0; st < S; st++) {
        const int *pbc = pbcs + 3 * st;
        const bool np = !pbc[0] || !pbc[1] || !pbc[2];
        memset(&geo[st], 0, sizeof(NLGeom));
        prepare_geom(lats + 9 * (size_t)st, pbc, cutoff, np && any_np ? mm.data() + 6 * st : nullptr, geo[st]);
        cbase[st] = (int)ncell;
        ncell += (long)geo[st].cells[0] * geo[st].cells[1] * geo[st].cells[2];
        if (ncell > (1L << 28)) return fail(GAPX_E_GEOMETRY, "too many cells (%ld)", ncell);
    }
    cbase[S] = (int)ncell;
    const int NC = (int)ncell;
    pl.NC = NC;
    // With lat_eff refreshed on the device (dev_geo), only the lattice-derived fields decide
    // whether the table is uploaded again; the last upload then came from an exact build or
    // an earlier speculative one of the same structures.
    const bool same = eg.have_geo && eg.h_geo.size() == geo.size() && eg.h_cell_base == cbase &&
                      std::equal(geo.begin(), geo.end(), eg.h_geo.begin(), dev_geo ? geo_equal_static : geo_equal);
    if (!same) {
        eg.have_geo = false;
        CHK(w.geo.ensure((size_t)S * sizeof(NLGeom) + 8));
        CHK(w.cell_base.ensure((size_t)(S + 1) * 4));
        const size_t gb = (size_t)S * sizeof(NLGeom), cb = (size_t)(S + 1) * 4;
        CHK(staging_ready(eg));
        CHK(eg.stage_geo.ensure(gb + cb + 64));
        char *h = eg.stage_geo.as<char>();
        memcpy(h, geo.data(), gb);
        memcpy(h + gb, cbase.data(), cb);
        HIPCHK(hipMemcpyAsync(w.geo.p, h, gb, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.cell_base.p, h + gb, cb, hipMemcpyHostToDevice, s));
        eg.staging_dirty = true;
        eg.h_geo = geo;
        eg.h_cell_base = cbase;
        eg.have_geo = true;
    }
    eg.ncell_last = NC;
    CHK(w.cell_of.ensure((size_t)N * 4 + 4));
    CHK(w.map_shift.ensure((size_t)N * 12 + 12));
    CHK(w.cell_cnt.ensure((size_t)NC * 4 + 4));
    CHK(w.cell_fill.ensure((size_t)NC * 4 + 4));
    CHK(w.cell_start.ensure((size_t)(NC + 1) * 4));
    CHK(w.cell_atoms.ensure((size_t)N * 4 + 4));
    CHK(w.cpos.ensure((size_t)N * sizeof(CellAtom) + sizeof(CellAtom)));
    CHK(w.cms.ensure((size_t)N * sizeof(CellShift) + sizeof(CellShift)));
    CHK(w.scan_part.ensure((size_t)scan_partials_size(std::max(NC, N) + 1) * 4 + 64));
    CHK(w.scan_part2.ensure((size_t)scan_partials_size(N + 1) * 4 + 64));
    CHK(w.n1_cnt.ensure((size_t)N * 4 + 4));
    CHK(w.n1_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n2_cnt.ensure((size_t)N * 4 + 4));
    CHK(w.n2_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n2_fill.ensure((size_t)N * 4 + 4));
    CHK(w.nb_off.ensure((size_t)(N + 1) * 4));
    CHK(w.cell_unsorted.ensure((size_t)N * 4 + 4));
    CHK(w.nl_flags.ensure(64));
    // The candidate search runs once, into per-atom slots sized from the largest counts
    // of the previous build; its counts are exact either way.  The first build counts
    // first, and a build whose slots overflowed fills the exact tables in a second
    // pass, as the first does.
    pl.cap1 = eg.nl_cap1;
    pl.cap2 = eg.nl_cap2;
    pl.one_pass = pl.cap1 > 0 && pl.cap2 > 0 && (size_t)N * std::max(pl.cap1, pl.cap2) < (size_t)INT32_MAX / 4;
    if (!pl.one_pass) spec = false;
    pl.spec = spec;
    pl.dev_geo = spec && dev_geo;
    if (pl.one_pass) {
        CHK(w.t1_j.ensure((size_t)N * pl.cap1 * 4 + 4));
        CHK(w.t1_shift.ensure((size_t)N * pl.cap1 * 12 + 12));
        CHK(w.t1_r.ensure((size_t)N * pl.cap1 * 8 + 8));
        CHK(w.t2_i.ensure((size_t)N * pl.cap2 * 4 + 4));
        CHK(w.t2_k.ensure((size_t)N * pl.cap2 * 4 + 4));
    }
    if (spec) {
        pl.capT1 = eg.nl_capT1;
        CHK(ensure_nl_tables(w, pl.capT1));
    }
    return GAPX_OK;
}

// The neighbour1/neighbour2 tables and the full view for up to T1 neighbour1 entries.
static int ensure_nl_tables(Workspace &w, int T1cap) {
    const size_t T1 = (size_t)T1cap + 1, P = 2 * (size_t)T1cap + 1;
    CHK(w.n1_j.ensure(T1 * 4));
    CHK(w.n1_shift.ensure(T1 * 12));
    CHK(w.n1_r.ensure(T1 * 8));
    CHK(w.n1_rev.ensure(T1 * 4));
    CHK(w.n2_i.ensure(T1 * 4));
    CHK(w.n2_k.ensure(T1 * 4));
    CHK(w.n2_is.ensure(T1 * 4));
    CHK(w.n2_ks.ensure(T1 * 4));
    CHK(w.nb_j.ensure(P * 4));
    CHK(w.nb_rev.ensure(P * 4));
    CHK(w.nb_r.ensure(P * 8));
    CHK(w.nb_d.ensure(P * 24));
    return GAPX_OK;
}

// The one-block front serves small problems whose geometry is on the host (the device
// geometry kernels write flags, which the front would clear).  Its search reads the
// candidates through the cells' atom lists; larger problems also write cell-ordered
// copies of positions and shifts, which their search reads without a dependent load
// (the copies cost the one-block front more than its search gains).
static bool nl_small(const NLPlan &pl) { return !pl.dev_geo && pl.N <= kNlFrontAtoms && pl.NC <= kNlFrontCells; }

// Binning, the candidate search and the count scans (flags: [0] overflow, [1] largest
// neighbour1 count, [2] largest neighbour2 count, [3] neighbour1 total).
static void nl_launch_front(Engine &eg, const NLPlan &pl, const double *d_pos) {
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int N = pl.N, NC = pl.NC;
    const NLGeom *dgeo = w.geo.as<NLGeom>();
    const int *dsid = w.sid.as<int>(), *dcb = w.cell_base.as<int>();
    int *flags = w.nl_flags.as<int>();
    const bool small = nl_small(pl);
    CellAtom *cpos = small ? nullptr : w.cpos.as<CellAtom>();
    CellShift *cms = small ? nullptr : w.cms.as<CellShift>();
    if (small) {
        launch_nl_front_small(N, NC, d_pos, dgeo, dsid, dcb, w.cell_of.as<int>(), w.map_shift.as<int>(),
                              w.cell_cnt.as<int>(), w.cell_start.as<int>(), w.cell_atoms.as<int>(), nullptr,
                              nullptr, w.n2_cnt.as<int>(), w.n2_fill.as<int>(), flags, s);
    } else {
        {
            int *zp[5] = {w.cell_cnt.as<int>(), w.cell_fill.as<int>(), w.n2_cnt.as<int>(), w.n2_fill.as<int>(),
                          flags};
            const int zn[5] = {NC, NC, N, N, 4};
            launch_zero_i32_multi(5, zp, zn, s);
        }
        if (pl.dev_geo) {
            launch_minmax_frac(pl.S, w.atom_off.as<int>(), d_pos, w.ginv.as<double>(), w.minmax.as<double>(), s);
            launch_geom_np(pl.S, w.atom_off.as<int>(), w.minmax.as<double>(), w.geo.as<NLGeom>(), flags, s);
        }
        launch_nl_bin(N, d_pos, dgeo, dsid, dcb, w.cell_of.as<int>(), w.map_shift.as<int>(), w.cell_cnt.as<int>(),
                      s);
        launch_scan_i32(w.cell_cnt.as<int>(), NC, w.cell_start.as<int>(), w.scan_part.as<int>(), s);
        launch_nl_cell_fill(N, w.cell_of.as<int>(), w.cell_start.as<int>(), w.cell_fill.as<int>(),
                            w.cell_unsorted.as<int>(), s);
        launch_nl_cell_rank(N, w.cell_of.as<int>(), w.cell_start.as<int>(), w.cell_cnt.as<int>(),
                            w.cell_unsorted.as<int>(), w.cell_atoms.as<int>(), d_pos, w.map_shift.as<int>(), cpos,
                            cms, s);
    }
    if (pl.one_pass) {
        launch_nl_pairs(2, N, d_pos, w.map_shift.as<int>(), w.cell_of.as<int>(), w.cell_start.as<int>(),
                        w.cell_cnt.as<int>(), w.cell_atoms.as<int>(), cpos, cms, dgeo, dsid, dcb,
                        w.n1_cnt.as<int>(), nullptr,
                        w.t1_j.as<int>(), w.t1_shift.as<int>(), w.t1_r.as<double>(), w.n2_cnt.as<int>(), nullptr,
                        nullptr, w.t2_i.as<int>(), w.t2_k.as<int>(), pl.cap1, pl.cap2, flags, s);
    } else {
        launch_nl_pairs(0, N, d_pos, w.map_shift.as<int>(), w.cell_of.as<int>(), w.cell_start.as<int>(),
                        w.cell_cnt.as<int>(), w.cell_atoms.as<int>(), cpos, cms, dgeo, dsid, dcb,
                        w.n1_cnt.as<int>(), nullptr,
                        nullptr, nullptr, nullptr, w.n2_cnt.as<int>(), nullptr, nullptr, nullptr, nullptr, 0, 0,
                        nullptr, s);
    }
    if (pl.spec && scan2_guard_fits(N)) {
        launch_scan2_guard(w.n1_cnt.as<int>(), w.n2_cnt.as<int>(), N, w.n1_off.as<int>(), w.n2_off.as<int>(), flags,
                           pl.capT1, s);
        return;
    }
    launch_scan2_max(w.n1_cnt.as<int>(), w.n2_cnt.as<int>(), N, w.n1_off.as<int>(), w.n2_off.as<int>(), flags + 1,
                     flags + 3, w.scan_part.as<int>(), w.scan_part2.as<int>(), s);
    if (pl.spec) launch_nl_guard(N, flags, pl.capT1, w.n1_off.as<int>(), w.n2_off.as<int>(), s);
}

// The tables from the slots (or a second, filling pass), neighbour2 in the reference's
// order and the full per-atom view.  slots_only (a speculative build, whose tables no
// later call reuses): the neighbour2 sort and the full view read the one-pass slots
// directly and the compacted neighbour1/neighbour2 copies are not made.
// slots_only with pairF: the full view's kernel also clears the step's pair forces,
// pair energies (pairE, if any) and site energies
static void nl_launch_back(Engine &eg, const NLPlan &pl, const double *d_pos, bool compact, bool slots_only = false,
                           double *pairF = nullptr, double *pairE = nullptr, double *esite = nullptr) {
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int N = pl.N;
    const NLGeom *dgeo = w.geo.as<NLGeom>();
    const int *dsid = w.sid.as<int>(), *dcb = w.cell_base.as<int>();
    if (compact && slots_only) {
        launch_nl_n2_sort(N, w.n2_off.as<int>(), w.t2_i.as<int>(), w.t2_k.as<int>(), w.cell_of.as<int>(),
                          w.n1_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(), w.n1_rev.as<int>(), pl.cap2, s);
        launch_nl_full(N, d_pos, dgeo, dsid, w.n1_off.as<int>(), w.t1_j.as<int>(), w.t1_shift.as<int>(),
                       w.t1_r.as<double>(), w.n2_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(),
                       w.n1_rev.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_rev.as<int>(),
                       w.nb_r.as<double>(), w.nb_d.as<double>(), pl.cap1, s, pairF, pairE, esite);
        return;
    }
    if (compact) {
        launch_nl_compact(N, pl.cap1, pl.cap2, w.n1_off.as<int>(), w.n2_off.as<int>(), w.t1_j.as<int>(),
                          w.t1_shift.as<int>(), w.t1_r.as<double>(), w.t2_i.as<int>(), w.t2_k.as<int>(),
                          w.n1_j.as<int>(), w.n1_shift.as<int>(), w.n1_r.as<double>(), w.n2_i.as<int>(),
                          w.n2_k.as<int>(), s);
    } else {
        launch_nl_pairs(1, N, d_pos, w.map_shift.as<int>(), w.cell_of.as<int>(), w.cell_start.as<int>(),
                        w.cell_cnt.as<int>(), w.cell_atoms.as<int>(), nl_small(pl) ? nullptr : w.cpos.as<CellAtom>(),
                        nl_small(pl) ? nullptr : w.cms.as<CellShift>(), dgeo, dsid, dcb, nullptr, w.n1_off.as<int>(),
                        w.n1_j.as<int>(), w.n1_shift.as<int>(), w.n1_r.as<double>(), nullptr, w.n2_off.as<int>(),
                        w.n2_fill.as<int>(), w.n2_i.as<int>(), w.n2_k.as<int>(), 0, 0, nullptr, s);
    }
    launch_nl_n2_sort(N, w.n2_off.as<int>(), w.n2_i.as<int>(), w.n2_k.as<int>(), w.cell_of.as<int>(),
                      w.n1_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(), w.n1_rev.as<int>(), 0, s);
    launch_nl_full(N, d_pos, dgeo, dsid, w.n1_off.as<int>(), w.n1_j.as<int>(), w.n1_shift.as<int>(),
                   w.n1_r.as<double>(), w.n2_off.as<int>(), w.n2_is.as<int>(), w.n2_ks.as<int>(), w.n1_rev.as<int>(),
                   w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_rev.as<int>(), w.nb_r.as<double>(),
                   w.nb_d.as<double>(), 0, s);
}

// next builds' capacities from a build's flags: per-atom slots of the largest counts plus a
// quarter and 8, the neighbour1 tables of the total plus an eighth and 1024 (grow-only)
static void update_nl_caps(Engine &eg, const int *flags) {
    eg.nl_cap1 = (flags[1] + flags[1] / 4 + 8 + 3) & ~3;
    eg.nl_cap2 = (flags[2] + flags[2] / 4 + 8 + 3) & ~3;
    const long t = (long)flags[3] + flags[3] / 8 + 1024;
    if (t < INT32_MAX / 4) eg.nl_capT1 = std::max(eg.nl_capT1, (int)t);
}

// The exact build: one host synchronisation for the neighbour1 total.
static int build_nl(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *d_pos, const double *h_pos,
                    const double *lats, const int *pbcs, double cutoff, int *P_out) {
    NLPlan pl;
    CHK(nl_prepare(m, eg, S, atom_off, d_pos, h_pos, lats, pbcs, cutoff, false, pl));
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    nl_launch_front(eg, pl, d_pos);
    CHK(eg.ctl.ensure(64));
    int *hcount = eg.ctl.as<int>();
    hcount[3] = -1;
    HIPCHK(hipMemcpyAsync(hcount, w.nl_flags.p, 16, hipMemcpyDeviceToHost, s));   // overflow, both maxima, total
    HIPCHK(hipStreamSynchronize(s));
    eg.staging_dirty = false;
    const int n1_total = hcount[3];
    const bool overflow = hcount[0] != 0;
    update_nl_caps(eg, hcount);
    if (eg.timer.on) eg.timer.resolve(eg.timer.cur ^ 1);   // the previous call has finished
    if (n1_total < 0 || n1_total > (INT32_MAX - 2) / 2) return fail(GAPX_E_GEOMETRY, "neighbour count overflow");
    CHK(ensure_nl_tables(w, n1_total));
    nl_launch_back(eg, pl, d_pos, pl.one_pass && !overflow);
    eg.n1_total_last = n1_total;
    *P_out = 2 * n1_total;
    return GAPX_OK;
}

// calc_connect with skin reuse (Connection.f95:1085-1127): with a cutoff above the
// model's (a skin), a list built for the same structures, cutoff and lattices is
// kept while no atom has moved by half the skin since that build; its distances
// are then recomputed as calc_dists does (Connection.f95:1321-1445) and the full
// per-atom view re-formed.  The displacement test and the distances run in one
// pass and are read back with one synchronisation, as the rebuild's count is.
static int connect(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *d_pos, const double *h_pos,
                   const double *lats, const int *pbcs, double cutoff, int *P_out) {
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int N = atom_off[S];
    const double skin = cutoff - m->cutoff;
    eg.last_rebuilt = true;
    if (skin > 0 && m->nl_reuse && eg.nl_valid && eg.nl_cut == cutoff && (int)eg.nl_pbcs.size() == 3 * S &&
        std::equal(lats, lats + 9 * S, eg.nl_lats.begin()) && std::equal(pbcs, pbcs + 3 * S, eg.nl_pbcs.begin())) {
        CHK(w.skin_flags.ensure(64));
        launch_fill_i32(w.skin_flags.as<int>(), 4, 0, s);
        unsigned long long *md = (unsigned long long *)w.skin_flags.p;
        int *ov = w.skin_flags.as<int>() + 2;
        launch_nl_skin_check(N, d_pos, w.last_pos.as<double>(), w.geo.as<NLGeom>(), w.sid.as<int>(),
                             w.n1_off.as<int>(), w.n1_j.as<int>(), w.n1_shift.as<int>(), w.n1_r.as<double>(), md, ov,
                             s);
        CHK(eg.ctl.ensure(64));
        HIPCHK(hipMemcpyAsync(eg.ctl.p, w.skin_flags.p, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        eg.staging_dirty = false;
        if (eg.timer.on) eg.timer.resolve(eg.timer.cur ^ 1);   // the previous call has finished
        double maxd;
        memcpy(&maxd, eg.ctl.p, 8);
        const int overlap = eg.ctl.as<int>()[2];
        if (maxd < 0.5 * skin) {
            if (overlap) return fail(GAPX_E_GEOMETRY, "calc_dists: two atoms overlap exactly");
            launch_nl_full(N, d_pos, w.geo.as<NLGeom>(), w.sid.as<int>(), w.n1_off.as<int>(), w.n1_j.as<int>(),
                           w.n1_shift.as<int>(), w.n1_r.as<double>(), w.n2_off.as<int>(), w.n2_is.as<int>(),
                           w.n2_ks.as<int>(), w.n1_rev.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                           w.nb_rev.as<int>(), w.nb_r.as<double>(), w.nb_d.as<double>(), 0, s);
            eg.last_rebuilt = false;
            *P_out = 2 * eg.n1_total_last;
            return GAPX_OK;
        }
    }
    eg.nl_valid = false;
    CHK(build_nl(m, eg, S, atom_off, d_pos, h_pos, lats, pbcs, cutoff, P_out));
    if (skin > 0 && m->nl_reuse) {   // remember this build for the next call
        CHK(w.last_pos.ensure((size_t)N * 24 + 24));
        if (N) HIPCHK(hipMemcpyAsync(w.last_pos.p, d_pos, (size_t)N * 24, hipMemcpyDeviceToDevice, s));
        eg.nl_cut = cutoff;
        eg.nl_lats.assign(lats, lats + 9 * S);
        eg.nl_pbcs.assign(pbcs, pbcs + 3 * S);
        eg.nl_valid = true;
    }
    return GAPX_OK;
}

// Scratch outputs.  Growing them frees the old blocks, so callers that hand
// these buffers in as outputs size them first.
static int ensure_outputs(Workspace &w, int S, int N) {
    CHK(w.f.ensure((size_t)N * 24 + 24));
    CHK(w.le.ensure((size_t)N * 8 + 8));
    CHK(w.lv.ensure((size_t)N * 72 + 72));
    CHK(w.e.ensure((size_t)S * 8 + 8));
    CHK(w.vir.ensure((size_t)S * 72 + 72));
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


int gapx::check_batch(int S, const int *atom_off, const int *Z, const double *lats, const int *pbcs) {
    if (S < 0) return fail(GAPX_E_ARG, "n_struct < 0");
    if (!atom_off) return fail(GAPX_E_ARG, "atom_off is NULL");
    if (atom_off[0] != 0) return fail(GAPX_E_ARG, "atom_off[0] must be 0");
    for (int st = 0; st < S; st++)
        if (atom_off[st + 1] < atom_off[st]) return fail(GAPX_E_ARG, "atom_off not ascending at %d", st);
    const int N = atom_off[S];
    if (N > 0 && !Z) return fail(GAPX_E_ARG, "Z is NULL");
    if (S > 0 && (!lats || !pbcs)) return fail(GAPX_E_ARG, "lattice / pbc are NULL");
    for (int i = 0; i < N; i++)
        if (Z[i] < 0 || Z[i] > 118) return fail(GAPX_E_ARG, "Z[%d]=%d out of range", i, Z[i]);
    return GAPX_OK;
}

// GEMM autotuner.  For each coordinate and centre-count bucket (ceil log2 nc) the
// first evaluation times every GEMM 1 tile and every GEMM 2 tile x split-K factor
// (fastest of three launches after a warm-up) with HIP events on the engine's stream, on the
// call's own operands (X is already computed), and keeps the fastest.  Later
// calls in the bucket reuse the choice.  Without autotuning: GEMM 1 (4,2),
// GEMM 2 (2,2) with the default split-K factor.
static int gemm_tune(gapx_model *m, Engine &eg, Coord &c, int nc, GemmTune &out, bool cached_only = false) {
    out = GemmTune();
    out.parts = gemm2_parts(nc, c.dim);
    const bool fusable = gp_f64_fused_supported(c.dim) && c.d_XsP.p;
    if (fusable) out.f64parts = gp_f64_parts_default(nc, c.M);   // the default: the fused kernel
    if (!m->autotune || nc <= 0) return GAPX_OK;
    int bucket = 0;
    while ((1 << bucket) < nc) bucket++;
    {
        std::lock_guard<std::mutex> lk(c.tune_mu);
        auto it = c.tune.find(bucket);
        if (it != c.tune.end()) {
            out = it->second;
            return GAPX_OK;
        }
    }
    if (cached_only) return GAPX_OK;   // the defaults: timing needs events
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    hipEvent_t ea, eb;
    HIPCHK(hipEventCreate(&ea));
    if (hipEventCreate(&eb) != hipSuccess) {
        hipEventDestroy(ea);
        return fail(GAPX_E_DEVICE, "hipEventCreate failed");
    }
    auto timed = [&](auto launch, float *ms) -> int {   // one warm-up, then the fastest of three launches
        launch();
        *ms = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            hipEventRecord(ea, s);
            launch();
            hipEventRecord(eb, s);
            HIPCHK(hipEventSynchronize(eb));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ea, eb));
            *ms = std::min(*ms, t);
        }
        return GAPX_OK;
    };
    GemmTune best;
    best.g1_ms = best.g2_ms = 1e30f;
    int st = GAPX_OK;
    for (int cfg = 0; cfg < kGemmTiles && st == GAPX_OK; cfg++) {
        float t = 0;
        int nt = 0;
        st = timed([&] { launch_gemm1_f64(cfg, nc, c.M, c.dim, w.X.as<double>(), c.d_Xs.as<double>(),
                                          c.d_w.as<double>(), c.zeta, w.W.as<double>(), w.e_part.as<double>(), &nt,
                                          s); }, &t);
        if (st == GAPX_OK && t < best.g1_ms) { best.g1_ms = t; best.g1 = cfg; }
    }
    for (int cfg = 0; cfg < kGemmTiles && st == GAPX_OK; cfg++)
        for (int parts = 1; parts <= 8 && st == GAPX_OK; parts *= 2) {
            float t = 0;
            st = timed([&] { launch_gemm2_f64(cfg, parts, nc, c.M, c.dim, w.W.as<double>(), c.d_Xs.as<double>(),
                                              w.G.as<double>(), w.Gparts.as<double>(), s); }, &t);
            if (st == GAPX_OK && t < best.g2_ms) { best.g2_ms = t; best.g2 = cfg; best.parts = parts; }
        }
    if (fusable) {   // the fused kernel (with its partial-G sum) against the two GEMMs
        float fbest = 1e30f;
        int fparts = 0;
        for (int p = 1, prev = 0; p <= 8 && st == GAPX_OK; p *= 2) {
            const int v = gp_f64_parts_valid(p, c.M);
            if (v == prev) continue;
            prev = v;
            float t = 0;
            int nt = 0;
            st = timed([&] {
                launch_gp_fused_f64(nc, c.M, c.dim, v, w.X.as<double>(), c.d_XsP.as<double>(), c.d_w.as<double>(),
                                    c.zeta, w.G.as<double>(), w.Gparts.as<double>(), w.e_part.as<double>(), &nt, s);
                launch_gp_finish_f64(nc, c.dim, nt, w.Gparts.as<double>(), w.G.as<double>(), nt, nullptr, nullptr,
                                     nullptr, s);
            }, &t);
            if (st == GAPX_OK && t < fbest) { fbest = t; fparts = v; }
        }
        best.f64_ms = fbest;
        best.f64parts = (st == GAPX_OK && fbest < best.g1_ms + best.g2_ms) ? fparts : 0;
    }
    hipEventDestroy(ea);
    hipEventDestroy(eb);
    CHK(st);
    HIPCHK(hipGetLastError());
    {
        std::lock_guard<std::mutex> lk(c.tune_mu);
        c.tune[bucket] = best;
    }
    out = best;
    return GAPX_OK;
}

// Split-form autotuner, as gemm_tune for the FP64 tiles: per coordinate, precision and
// centre-count bucket the first evaluation times the fused kernel (with its partial-G sum)
// for 1, 2, 4 and 8 sparse-point parts and the stream-K schedule (-1), or the split GEMMs
// with 64 x 64 and 128 x 128 tiles, on the call's own operands, and keeps the fastest
// (one warm-up, best of three).
// Bytes of the partial-G buffer the fused kernel needs with `parts` (0: the default rule)
static size_t fused_scratch_bytes(int prec, const Coord &c, int nc, int parts) {
    const int dpad = (c.dim + 31) / 32 * 32, Mpad = (c.M + 31) / 32 * 32;
    if (parts < 0) return gp_fused_sk_scratch_bytes(prec, c.dim, dpad, nc, Mpad);
    const int v = parts > 0 ? gp_fused_parts_valid(parts, Mpad) : gp_fused_parts(nc, Mpad);
    return v > 1 ? (size_t)v * nc * c.dim * 4 : 0;
}
static int fused_part_candidates(int prec, const Coord &c, int nc, std::vector<int> &out) {
    out.clear();
    const int Mpad = (c.M + 31) / 32 * 32;
    for (int p = 1; p <= 8; p *= 2) {
        const int v = gp_fused_parts_valid(p, Mpad);
        if ((v == 1 || (size_t)v * nc * c.dim * 4 <= kTuneScratchCap) &&
            std::find(out.begin(), out.end(), v) == out.end())
            out.push_back(v);
    }
    if (fused_scratch_bytes(prec, c, nc, -1) <= kTuneScratchCap) out.push_back(-1);
    return GAPX_OK;
}

static int mixed_tune(gapx_model *m, Engine &eg, Coord &c, int nc, bool fused, GemmTune &out, bool cached_only) {
    out = GemmTune();
    if (!m->autotune || nc <= 0) return GAPX_OK;
    int bucket = 0;
    while ((1 << bucket) < nc) bucket++;
    const int key = (m->precision << 8) | bucket;
    {
        std::lock_guard<std::mutex> lk(c.tune_mu);
        auto it = c.mtune.find(key);
        if (it != c.mtune.end()) {
            out = it->second;
            return GAPX_OK;
        }
    }
    if (cached_only) return GAPX_OK;
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    hipEvent_t ea, eb;
    HIPCHK(hipEventCreate(&ea));
    if (hipEventCreate(&eb) != hipSuccess) {
        hipEventDestroy(ea);
        return fail(GAPX_E_DEVICE, "hipEventCreate failed");
    }
    auto timed = [&](auto launch, float *ms) -> int {
        launch();
        *ms = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            hipEventRecord(ea, s);
            launch();
            hipEventRecord(eb, s);
            HIPCHK(hipEventSynchronize(eb));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ea, eb));
            *ms = std::min(*ms, t);
        }
        return GAPX_OK;
    };
    GemmTune best;
    best.g1_ms = best.g2_ms = 1e30f;
    int st = GAPX_OK;
    if (fused) {
        std::vector<int> cand;
        fused_part_candidates(m->precision, c, nc, cand);
        for (int parts : cand) {
            float t = 0;
            int nt = 0;
            st = timed([&] {
                launch_gp_fused_split(m->precision, nc, c.M, c.dim, c.dpad, c.Mpad, parts, w.X.as<double>(),
                                      c.d_Xs_hi.p, c.d_Xs_lo.p, c.d_w.as<double>(), c.zeta, c.wscale,
                                      w.G.as<double>(), w.Gparts.as<float>(), w.e_part.as<double>(), &nt, s);
                if (nt < 0)
                    launch_gp_sk_finish(m->precision, nc, c.dim, c.dpad, c.Mpad, true, w.Gparts.as<float>(), c.wscale,
                                        w.G.as<double>(), w.e_part.as<double>(), nullptr, nullptr, s);
                else
                    launch_gp_fused_reduce(nc, c.dim, nt, w.Gparts.as<float>(), c.wscale, w.G.as<double>(), s);
            }, &t);
            if (st != GAPX_OK) break;
            if (t < best.g1_ms) { best.g1_ms = t; best.fparts = parts; }
        }
        best.g2_ms = 0;
    } else {
        for (int lds = 0; lds <= 1 && st == GAPX_OK; lds++) {
            float t = 0;
            int nt = 0;
            st = timed([&] { launch_gemm1_split(m->precision, nc, c.M, c.dim, c.dpad, lds, w.X.as<double>(),
                                                c.d_Xs_hi.p, c.d_Xs_lo.p, c.d_w.as<double>(), c.zeta, c.wscale,
                                                w.Wf.as<float>(), w.e_part.as<double>(), &nt, w.split_ws.p, s); }, &t);
            if (st == GAPX_OK && t < best.g1_ms) { best.g1_ms = t; best.g1lds = lds; }
        }
        for (int lds = 0; lds <= 1 && st == GAPX_OK; lds++) {
            float t = 0;
            st = timed([&] { launch_gemm2_split(m->precision, nc, c.M, c.dim, c.dpad, c.Mpad, lds, w.Wf.as<float>(),
                                                c.d_XsT_hi.p, c.d_XsT_lo.p, c.wscale, w.G.as<double>(), s); }, &t);
            if (st == GAPX_OK && t < best.g2_ms) { best.g2_ms = t; best.g2lds = lds; }
        }
    }
    hipEventDestroy(ea);
    hipEventDestroy(eb);
    CHK(st);
    HIPCHK(hipGetLastError());
    {
        std::lock_guard<std::mutex> lk(c.tune_mu);
        c.mtune[key] = best;
    }
    out = best;
    return GAPX_OK;
}

extern "C" int gapx_model_tuning_split(gapx_model *m, int coord, int max_entries, int *precision, int *bucket,
                                       int *fused_parts, int *gemm1_lds, int *gemm2_lds, double *fused_or_gemm1_ms,
                                       double *gemm2_ms, int *n_entries) {
    if (!m || !n_entries) return fail(GAPX_E_ARG, "NULL argument");
    if (coord < 0 || coord >= (int)m->coords.size()) return fail(GAPX_E_ARG, "coordinate %d out of range", coord);
    Coord &c = *m->coords[coord];
    std::lock_guard<std::mutex> lk(c.tune_mu);
    int k = 0;
    for (auto &e : c.mtune) {
        if (k < max_entries) {
            if (precision) precision[k] = e.first >> 8;
            if (bucket) bucket[k] = e.first & 255;
            if (fused_parts) fused_parts[k] = e.second.fparts;
            if (gemm1_lds) gemm1_lds[k] = e.second.g1lds;
            if (gemm2_lds) gemm2_lds[k] = e.second.g2lds;
            if (fused_or_gemm1_ms) fused_or_gemm1_ms[k] = e.second.g1_ms;
            if (gemm2_ms) gemm2_ms[k] = e.second.g2_ms;
        }
        k++;
    }
    *n_entries = k;
    return GAPX_OK;
}

// GAP variance set-up (gpCoordinates_initialise_variance_estimate's role, IPModel_GAP.f95:412-414):
// Cinv = (K_MM + reg I)^-1 with K_MM[s][t] = delta^2 kernel(X_s, X_t), formed once per
// coordinate and regularisation on the host (Cholesky, triangular inverse, Linv^T Linv;
// rows split over host threads) and kept on the device.
template <class F> static void parallel_rows(int n, F fn) {
    const int nt = std::max(1, std::min<int>((int)std::thread::hardware_concurrency(), n / 64 + 1));
    if (nt <= 1) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
            for (int i = next++; i < n; i = next++) fn(i);
        });
    for (auto &x : th) x.join();
}

static int var_prepare(Coord &c, double reg, hipStream_t s) {
    std::lock_guard<std::mutex> lk(c.var_mu);
    if (c.var_reg == reg) return GAPX_OK;
    if (!(reg >= 0)) return fail(GAPX_E_ARG, "gap_variance_regularisation must be >= 0 (got %g)", reg);
    for (double sc : c.scut)
        if (sc != 1.0)
            return fail(GAPX_E_UNSUPPORTED, "gpCoordinates %s: the GAP variance with sparseCutoff != 1 is not built",
                        c.label.c_str());
    if (c.kind == 2 && c.M > kMaxVarSparse)
        return fail(GAPX_E_UNSUPPORTED, "distance_2b variance with %d sparse points (at most %d)", c.M, kMaxVarSparse);
    const int M = c.M, d = c.dim;
    const double d2 = c.delta * c.delta;
    std::vector<double> L((size_t)M * M, 0.0), LT((size_t)M * M, 0.0), C((size_t)M * M, 0.0);
    parallel_rows(M, [&](int a) {   // K_MM + reg I, lower triangle
        const double *xa = c.Xs.data() + (size_t)a * d;
        for (int b = 0; b <= a; b++) {
            const double *xb = c.Xs.data() + (size_t)b * d;
            double k;
            if (c.kind == 1) {
                double dot = 0.0;
                for (int q = 0; q < d; q++) dot += xa[q] * xb[q];
                k = d2 * std::pow(dot, c.zeta);
            } else {
                const double t = (xa[0] - xb[0]) / c.theta[0];
                k = d2 * std::exp(-0.5 * t * t);
            }
            L[(size_t)a * M + b] = k + (a == b ? reg : 0.0);
        }
    });
    for (int i = 0; i < M; i++) {   // Cholesky-Crout in place: row i from rows 0..i-1
        double *li = L.data() + (size_t)i * M;
        for (int j = 0; j <= i; j++) {
            const double *lj = L.data() + (size_t)j * M;
            double t = li[j];
            for (int k = 0; k < j; k++) t -= li[k] * lj[k];
            if (j < i) li[j] = t / lj[j];
            else {
                if (!(t > 0))
                    return fail(GAPX_E_MODEL, "gpCoordinates %s: K_MM + %g I is not positive definite", c.label.c_str(),
                                reg);
                li[i] = std::sqrt(t);
            }
        }
    }
    parallel_rows(M, [&](int col) {   // LT[col][k] = (L^-1)[k][col], k >= col
        double *t = LT.data() + (size_t)col * M;
        t[col] = 1.0 / L[(size_t)col * M + col];
        for (int i = col + 1; i < M; i++) {
            const double *li = L.data() + (size_t)i * M;
            double a = 0.0;
            for (int k = col; k < i; k++) a += li[k] * t[k];
            t[i] = -a / li[i];
        }
    });
    parallel_rows(M, [&](int a) {   // C = L^-T L^-1: C[a][b] = sum_k LT[a][k] LT[b][k], k >= max(a, b)
        const double *ta = LT.data() + (size_t)a * M;
        for (int b = 0; b <= a; b++) {
            const double *tb = LT.data() + (size_t)b * M;
            double v = 0.0;
            for (int k = a; k < M; k++) v += ta[k] * tb[k];
            C[(size_t)a * M + b] = v;
        }
    });
    for (int a = 0; a < M; a++)
        for (int b = a + 1; b < M; b++) C[(size_t)a * M + b] = C[(size_t)b * M + a];
    CHK(c.d_Cinv.ensure((size_t)M * M * 8 + 8));
    HIPCHK(hipMemcpyAsync(c.d_Cinv.p, C.data(), (size_t)M * M * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c.var_reg = reg;
    return GAPX_OK;
}

// Device-resident evaluation of S structures in one pass (IPModel_GAP_Calc per
// structure).  Outputs: d_e[S], d_v[9S] and the per-atom arrays; any may be NULL,
// in which case the engine's scratch takes it.
// The connection of a LAMMPS pair style (quip_lammps_wrapper.f95:104-144):
// neighbour1/neighbour2 built on the host from LAMMPS's lists, all shifts zero
// (periodic images are explicit ghost atoms).  The per-atom full view is formed
// on the GPU by the kernel calc_connect uses.  The list is not kept for skin
// reuse.
static int connect_external(Engine &eg, const ExtNL &x, int N, const double *d_pos, const double *lat, int *P_out) {
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int n1_total = x.n1_off[N];
    const size_t T1 = (size_t)n1_total + 1, P = 2 * (size_t)n1_total + 1;
    eg.nl_valid = false;
    eg.have_geo = false;
    NLGeom g;
    memset(&g, 0, sizeof(NLGeom));
    for (int k = 0; k < 9; k++) g.lat[k] = lat[k];
    CHK(w.geo.ensure(sizeof(NLGeom) + 8));
    CHK(w.n1_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n2_off.ensure((size_t)(N + 1) * 4));
    CHK(w.nb_off.ensure((size_t)(N + 1) * 4));
    CHK(w.n1_j.ensure(T1 * 4));
    CHK(w.n1_shift.ensure(T1 * 12));
    CHK(w.n1_r.ensure(T1 * 8));
    CHK(w.n1_rev.ensure(T1 * 4));
    CHK(w.n2_is.ensure(T1 * 4));
    CHK(w.n2_ks.ensure(T1 * 4));
    CHK(w.nb_j.ensure(P * 4));
    CHK(w.nb_rev.ensure(P * 4));
    CHK(w.nb_r.ensure(P * 8));
    CHK(w.nb_d.ensure(P * 24));
    // pageable sources, read before this function returns (synchronised below)
    HIPCHK(hipMemcpyAsync(w.geo.p, &g, sizeof(NLGeom), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.n1_off.p, x.n1_off.data(), (size_t)(N + 1) * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.n2_off.p, x.n2_off.data(), (size_t)(N + 1) * 4, hipMemcpyHostToDevice, s));
    if (n1_total > 0) {
        HIPCHK(hipMemcpyAsync(w.n1_j.p, x.n1_j.data(), (size_t)n1_total * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.n1_r.p, x.n1_r.data(), (size_t)n1_total * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.n1_rev.p, x.n1_rev.data(), (size_t)n1_total * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.n2_is.p, x.n2_i.data(), (size_t)n1_total * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.n2_ks.p, x.n2_k.data(), (size_t)n1_total * 4, hipMemcpyHostToDevice, s));
        launch_fill_i32(w.n1_shift.as<int>(), 3 * n1_total, 0, s);
    }
    HIPCHK(hipStreamSynchronize(s));
    launch_nl_full(N, d_pos, w.geo.as<NLGeom>(), w.sid.as<int>(), w.n1_off.as<int>(), w.n1_j.as<int>(),
                   w.n1_shift.as<int>(), w.n1_r.as<double>(), w.n2_off.as<int>(), w.n2_is.as<int>(),
                   w.n2_ks.as<int>(), w.n1_rev.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_rev.as<int>(),
                   w.nb_r.as<double>(), w.nb_d.as<double>(), 0, s);
    eg.n1_total_last = n1_total;
    *P_out = (int)(P - 1);
    return GAPX_OK;
}

static int resolve_pending_list(gapx_model *m, Engine &eg, std::vector<Engine::Pending> *);

int gapx::eval_batch(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *d_pos, const double *h_pos,
                     const int *Z, const double *lats, const int *pbcs, double nl_cutoff, unsigned want, double *d_e,
                     double *d_le, double *d_f, double *d_v, double *d_lv, const ExtNL *ext, const CalcOpts *opt,
                     bool noalloc) {
    CHK(check_batch(S, atom_off, Z, lats, pbcs));
    const int N = atom_off[S];
    if (N > 0 && !d_pos) return fail(GAPX_E_ARG, "positions are NULL");
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    StageTimer &tm = eg.timer;
    HIPCHK(hipSetDevice(m->device));
    const double rc = nl_cutoff > 0 ? nl_cutoff : m->cutoff;
    if (rc < m->cutoff) return fail(GAPX_E_ARG, "nl_cutoff %g is below the model cutoff %g", rc, m->cutoff);
    const bool grad = (want & (GAPX_WANT_F | GAPX_WANT_V | GAPX_WANT_LOCAL_V)) != 0;
    if (ext && S != 1) return fail(GAPX_E_ARG, "a supplied connection covers one structure");
    if (ext && opt) return fail(GAPX_E_ARG, "calc options with a supplied connection");
    if (opt && opt->only_coord < 0) return fail(GAPX_E_ARG, "only_descriptor must be >= 0");
    const bool want_var = opt && opt->variance;
    const bool want_vgrad = want_var && opt->d_gvg;
    // The speculative neighbour list: no read-back inside the call, so the whole call is
    // one launch sequence (and graph).  Not for the first build, a skin (whose reuse test
    // reads back), a supplied connection, calc options or the allocation-free entry.
    const bool spec_ok = !ext && !opt && !noalloc && !eg.force_exact && !(rc > m->cutoff && m->nl_reuse);
    // Calls queued speculatively are read (and redone if they overflowed) before anything
    // of this call is staged or queued: before a call that is not speculative, and when
    // the read-back area is full.
    if (!eg.force_exact) {
        const size_t need_max = (16 + 48 * (size_t)S + 15) & ~(size_t)15;
        if (!spec_ok || eg.pending.size() >= 64 || eg.checks_used + need_max > eg.checks.cap) {
            CHK(resolve_pending_list(m, eg, nullptr));
            eg.checks.coherent = true;   // written by the step's last kernel in eager calls
            if (spec_ok) CHK(eg.checks.ensure(std::max<size_t>(need_max * 64, 4096)));
        }
    }
    if (tm.on) tm.begin(s);
    const unsigned char *hmask = ext ? ext->local.data() : (opt ? opt->mask : nullptr);
    CHK(stage_inputs(m, eg, S, atom_off, Z, hmask));
    const unsigned char *cmask = hmask ? w.cmask.as<unsigned char>() : nullptr;
    // only_descriptor: with several coordinates, coordinate k alone (IPModel_GAP.f95:398-403)
    const size_t NC0 = m->coords.size();
    auto skipped = [&](size_t k) { return opt && opt->only_coord > 0 && NC0 > 1 && (int)k != opt->only_coord - 1; };
    NLPlan pl;
    if (spec_ok) CHK(nl_prepare(m, eg, S, atom_off, d_pos, h_pos, lats, pbcs, rc, true, pl));
    const bool spec = spec_ok && pl.spec;
    const int t_nl0 = tm.on ? tm.mark(s) : -1;
    int P = 0;
    if (spec) {
        P = 2 * pl.capT1;   // a capacity: kernels that loop over pairs read their counts on the device
        eg.nl_valid = false;
        eg.last_rebuilt = true;
    } else if (spec_ok) {   // prepared, but without capacities yet: the exact build from here
        nl_launch_front(eg, pl, d_pos);
        CHK(eg.ctl.ensure(64));
        int *hc = eg.ctl.as<int>();
        HIPCHK(hipMemcpyAsync(hc, w.nl_flags.p, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        eg.staging_dirty = false;
        update_nl_caps(eg, hc);
        if (eg.timer.on) eg.timer.resolve(eg.timer.cur ^ 1);
        const int n1_total = hc[3];
        if (n1_total < 0 || n1_total > (INT32_MAX - 2) / 2) return fail(GAPX_E_GEOMETRY, "neighbour count overflow");
        CHK(ensure_nl_tables(w, n1_total));
        nl_launch_back(eg, pl, d_pos, pl.one_pass && hc[0] == 0);
        eg.n1_total_last = n1_total;
        eg.nl_valid = false;
        eg.last_rebuilt = true;
        P = 2 * n1_total;
    } else if (ext) CHK(connect_external(eg, *ext, N, d_pos, lats, &P));
    else CHK(connect(m, eg, S, atom_off, d_pos, h_pos, lats, pbcs, rc, &P));
    if (tm.on && !spec) tm.stage(0, t_nl0, tm.mark(s));

    // Prepare: every buffer the launches use is sized here (to the largest
    // coordinate), tiles are chosen and split copies made, so that the launch
    // sequence below allocates nothing and can be captured as a graph.
    const size_t NCOORD = m->coords.size();
    std::vector<GemmTune> tunes(NCOORD);
    std::vector<char> fused(NCOORD, 0);
    std::vector<size_t> gp_bytes(NCOORD, 0);   // partial-G bytes sized per coordinate (split fused kernel)
    size_t mX = 8, mN = 8, mC = 8, mW = 8, mE = 8, mGp = 8, mWf = 8, mSw = 8, mV = 8;
    bool has_d2b_var = false;
    for (size_t k = 0; k < NCOORD; k++) {
        Coord &c = *m->coords[k];
        const int nc = (int)eg.h_centres[k].size();
        if (skipped(k)) continue;
        if (want_var) {
            CHK(var_prepare(c, opt->reg, s));
            if (c.kind == 2) {
                has_d2b_var = true;
                CHK(w.var_inst[k].ensure((size_t)(P + 1) * 8));
                CHK(w.cut2_inst[k].ensure((size_t)(P + 1) * 8));
            } else {
                CHK(w.var_inst[k].ensure((size_t)(nc + 1) * 8));
                mV = std::max(mV, (size_t)nc * c.M * 8);
            }
        }
        if (c.kind != 1 || nc == 0) continue;
        const size_t KLM = (size_t)c.S * c.n * (c.L + 1) * (c.L + 1);
        const int dpad = (c.dim + 31) / 32 * 32, Mpad = (c.M + 31) / 32 * 32;
        fused[k] = m->precision != GAPX_PREC_FP64 && gp_fused_supported(dpad, Mpad);
        mX = std::max(mX, (size_t)nc * c.dim * 8);
        mN = std::max(mN, (size_t)nc * 8);
        mC = std::max(mC, (size_t)nc * KLM * 8);
        mE = std::max(mE, (size_t)nc * ((c.M + 31) / 32 + 1) * 8);
        if (m->precision == GAPX_PREC_FP64) {
            mW = std::max(mW, (size_t)nc * c.M * 8);
            mGp = std::max(mGp, (size_t)8 * nc * c.dim * 8);
        } else if (fused[k]) {
            size_t b = fused_scratch_bytes(m->precision, c, nc, 0);
            if (m->autotune && !noalloc) {   // the tuner's candidates
                std::vector<int> cand;
                fused_part_candidates(m->precision, c, nc, cand);
                for (int p : cand) b = std::max(b, fused_scratch_bytes(m->precision, c, nc, p));
            } else if (m->autotune) {   // allocation-free: a tuning cached by an allocating call, if its slots are modest
                GemmTune cached;
                CHK(mixed_tune(m, eg, c, nc, true, cached, true));
                const size_t cb = fused_scratch_bytes(m->precision, c, nc, cached.fparts);
                if (cb <= kNoallocTuneCap) b = std::max(b, cb);
            }
            if (m->fused_schedule) b = std::max(b, fused_scratch_bytes(m->precision, c, nc, m->fused_schedule));
            gp_bytes[k] = b;
            mGp = std::max(mGp, b);
        } else {
            mWf = std::max(mWf, (size_t)nc * Mpad * 4);
            mSw = std::max(mSw, (size_t)(nc + 63) / 64 * 64 * dpad * 4);
        }
    }
    CHK(w.X.ensure(mX)); CHK(w.G.ensure(mX)); CHK(w.norm.ensure(mN)); CHK(w.cbuf.ensure(mC));
    CHK(w.e_part.ensure(mE)); CHK(w.W.ensure(mW)); CHK(w.Gparts.ensure(mGp + 16));
    CHK(w.Wf.ensure(mWf + 64)); CHK(w.split_ws.ensure(mSw + 64));
    CHK(w.pairF.ensure((size_t)(P + 1) * 24));
    CHK(w.pairE.ensure((size_t)(P + 1) * 8));
    CHK(w.esite.ensure((size_t)(N + 1) * 8));
    if (opt && opt->d_epc) CHK(w.eatom.ensure((size_t)(N + 1) * 8));
    if (want_var) {
        CHK(w.vK.ensure(mV)); CHK(w.vk.ensure(mV)); CHK(w.vV.ensure(mV));
        CHK(w.Gv.ensure(mX));
        CHK(w.pairFv.ensure((size_t)(P + 1) * 24));
        CHK(w.pairVE.ensure((size_t)(P + 1) * 8));
        CHK(w.varsite.ensure((size_t)(N + 1) * 8));
    }
    CHK(ensure_outputs(w, S, N));
    for (size_t k = 0; k < NCOORD; k++) {
        Coord &c = *m->coords[k];
        const int nc = (int)eg.h_centres[k].size();
        if (c.kind != 1 || nc == 0 || skipped(k)) continue;
        if (m->precision == GAPX_PREC_FP64) CHK(gemm_tune(m, eg, c, nc, tunes[k], noalloc));
        else {
            if (c.split_prec != m->precision) {
                if (noalloc)
                    return fail(GAPX_E_ARG, "split copies of the sparse points are missing (set the precision first)");
                CHK(ensure_split(m, c, m->precision, s));
            }
            CHK(mixed_tune(m, eg, c, nc, fused[k], tunes[k], noalloc));
            if (fused[k] && m->fused_schedule) tunes[k].fparts = m->fused_schedule;
            // a tuning cached by an allocating call may need more partial-G space than the
            // allocation-free call sized (over kNoallocTuneCap): the default rule then
            if (fused[k] && fused_scratch_bytes(m->precision, c, nc, tunes[k].fparts) > gp_bytes[k])
                tunes[k].fparts = 0;
        }
    }
    double *o_e = (want & GAPX_WANT_E) && d_e ? d_e : w.e.as<double>();
    double *o_v = (want & GAPX_WANT_V) && d_v ? d_v : w.vir.as<double>();
    double *o_f = (want & GAPX_WANT_F) && d_f ? d_f : w.f.as<double>();
    double *o_le = (want & GAPX_WANT_LOCAL_E) && d_le ? d_le : w.le.as<double>();
    double *o_lv = (want & GAPX_WANT_LOCAL_V) && d_lv ? d_lv : w.lv.as<double>();

    // The launch sequence (stage events only when timing, never inside a graph).
    // pair energies exist only with a distance_2b term; without one the gather skips them
    bool has_d2b = false;
    for (size_t k = 0; k < NCOORD; k++) has_d2b = has_d2b || m->coords[k]->kind == 2;
    double *pairE = has_d2b ? w.pairE.as<double>() : nullptr;
    double *d_epc = opt ? opt->d_epc : nullptr;
    if (S == 1) CHK(w.spart.ensure((size_t)(N / 8 + 2) * 80));   // the gather's block partials
    // Eager speculative calls hand the build flags to their read-back slot from the step's
    // last kernel (the slot is this call's: pending calls take consecutive slots); a graph
    // replay keeps a copy after the graph, whose kernel arguments are fixed at capture.
    const bool use_graph = m->graphs && !tm.on && !ext && !opt && !noalloc;
    int *flag_slot = spec && !use_graph ? (int *)(eg.checks.as<char>() + eg.checks_used) : nullptr;
    auto launch_all = [&](bool marks) {
        if (spec) {   // the neighbour list joins the sequence; pair arrays are cleared to its count
            nl_launch_front(eg, pl, d_pos);
            nl_launch_back(eg, pl, d_pos, true, true, w.pairF.as<double>(), pairE, w.esite.as<double>());
            if (marks) tm.stage(0, t_nl0, tm.mark(s));
        } else {
            launch_zero_step(w.pairF.as<double>(), pairE, w.esite.as<double>(), N, nullptr, P, (size_t)P, s);
        }
        if (d_epc) launch_fill_f64(d_epc, NCOORD, 0.0, s);
        if (want_var) {
            double *zp[3] = {w.pairFv.as<double>(), w.varsite.as<double>(), w.pairVE.as<double>()};
            const size_t zn[3] = {(size_t)P * 3, (size_t)N, (size_t)P};
            launch_zero_f64_multi(has_d2b_var ? 3 : 2, zp, zn, s);
        }
        for (size_t k = 0; k < NCOORD; k++) {
            Coord &c = *m->coords[k];
            if (skipped(k)) continue;
            const double d2 = c.delta * c.delta;
            if (c.kind == 2) {
                const int ta = marks ? tm.mark(s) : -1;
                launch_d2b(N, w.Z.as<int>(), c.Z1, c.Z2, c.rc, c.ctw, c.M, c.d_Xs.as<double>(), c.d_w.as<double>(),
                           c.d_theta.as<double>(), w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_r.as<double>(),
                           w.nb_d.as<double>(), w.pairF.as<double>(), pairE, cmask,
                           d_epc ? w.eatom.as<double>() : nullptr, s);
                if (d_epc) launch_sum_fixed(N, w.eatom.as<double>(), d_epc + k, s);
                if (want_var) {
                    const double nan = std::nan("");
                    launch_fill_f64(w.var_inst[k].as<double>(), P, nan, s);
                    launch_fill_f64(w.cut2_inst[k].as<double>(), P, nan, s);
                    launch_d2b_var(N, w.Z.as<int>(), c.Z1, c.Z2, c.rc, c.ctw, c.M, c.d_Xs.as<double>(), d2,
                                   c.d_theta.as<double>(), c.d_Cinv.as<double>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                                   w.nb_r.as<double>(), w.nb_d.as<double>(), w.pairFv.as<double>(),
                                   w.pairVE.as<double>(), w.var_inst[k].as<double>(), w.cut2_inst[k].as<double>(),
                                   cmask, s);
                }
                if (marks) tm.stage(5, ta, tm.mark(s));
                continue;
            }
            const int nc = (int)eg.h_centres[k].size();
            if (nc == 0) continue;
            const int *centres = w.centres[k].as<int>();
            const GemmTune &tu = tunes[k];
            SoapDev sp = soap_dev(c);
            const int t0 = marks ? tm.mark(s) : -1;
            launch_soap_fwd(sp, nc, centres, w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(), w.nb_r.as<double>(),
                            w.nb_d.as<double>(), w.X.as<double>(), w.norm.as<double>(), w.cbuf.as<double>(), s);
            const int t1 = marks ? tm.mark(s) : -1;
            int ntile = 0;
            const bool f64fused = m->precision == GAPX_PREC_FP64 && grad && tu.f64parts > 0;
            if (f64fused) {   // K -> W -> G in one kernel
                launch_gp_fused_f64(nc, c.M, c.dim, tu.f64parts, w.X.as<double>(), c.d_XsP.as<double>(),
                                    c.d_w.as<double>(), c.zeta, w.G.as<double>(), w.Gparts.as<double>(),
                                    w.e_part.as<double>(), &ntile, s);
            } else if (m->precision == GAPX_PREC_FP64) {
                launch_gemm1_f64(tu.g1, nc, c.M, c.dim, w.X.as<double>(), c.d_Xs.as<double>(), c.d_w.as<double>(),
                                 c.zeta, w.W.as<double>(), w.e_part.as<double>(), &ntile, s);
            } else if (fused[k]) {   // split forms: K -> W -> G in one kernel
                launch_gp_fused_split(m->precision, nc, c.M, c.dim, c.dpad, c.Mpad, tu.fparts, w.X.as<double>(),
                                      c.d_Xs_hi.p, c.d_Xs_lo.p, c.d_w.as<double>(), c.zeta, c.wscale,
                                      w.G.as<double>(), w.Gparts.as<float>(), w.e_part.as<double>(), &ntile, s);
            } else {
                launch_gemm1_split(m->precision, nc, c.M, c.dim, c.dpad, tu.g1lds, w.X.as<double>(), c.d_Xs_hi.p,
                                   c.d_Xs_lo.p,
                                   c.d_w.as<double>(), c.zeta, c.wscale, w.Wf.as<float>(), w.e_part.as<double>(),
                                   &ntile, w.split_ws.p, s);
            }
            const int t2 = marks ? tm.mark(s) : -1;
            // the energy rows, with the partial-G sum of a fused kernel in the same launch
            const bool split_fused = m->precision != GAPX_PREC_FP64 && fused[k];
            if (f64fused)
                launch_gp_finish_f64(nc, c.dim, ntile, w.Gparts.as<double>(), w.G.as<double>(), ntile, centres,
                                     w.e_part.as<double>(), w.esite.as<double>(), s);
            else if (split_fused && ntile < 0) {   // stream-K schedule: shared row blocks and every energy
                launch_gp_sk_finish(m->precision, nc, c.dim, c.dpad, c.Mpad, grad, w.Gparts.as<float>(), c.wscale,
                                    w.G.as<double>(), w.e_part.as<double>(), centres, w.esite.as<double>(), s);
                ntile = 1;
            } else if (split_fused && grad)
                launch_gp_finish_f32(nc, c.dim, ntile, w.Gparts.as<float>(), 1.0 / c.wscale, w.G.as<double>(), ntile,
                                     centres, w.e_part.as<double>(), w.esite.as<double>(), s);
            else
                launch_energy_rows(nc, ntile, centres, w.e_part.as<double>(), w.esite.as<double>(), s);
            if (d_epc) launch_sum_fixed((size_t)nc * ntile, w.e_part.as<double>(), d_epc + k, s);
            if (marks) { tm.stage(1, t0, t1); tm.stage(2, t1, t2); }
            if (want_var) {   // FP64 whatever the GEMM precision
                launch_gemm_nt_f64(nc, c.M, c.dim, w.X.as<double>(), c.d_Xs.as<double>(), w.vK.as<double>(), s);
                launch_var_cov(nc, c.M, w.vK.as<double>(), d2, c.zeta, w.vk.as<double>(), s);
                launch_gemm2_f64(1, 1, nc, c.M, c.M, w.vk.as<double>(), c.d_Cinv.as<double>(), w.vV.as<double>(),
                                 nullptr, s);
                launch_var_rows(nc, c.M, c.dim, w.X.as<double>(), w.vK.as<double>(), w.vk.as<double>(),
                                w.vV.as<double>(), d2, c.zeta, centres, w.var_inst[k].as<double>(),
                                w.varsite.as<double>(), s);
                if (want_vgrad) {
                    launch_gemm2_f64(1, 1, nc, c.M, c.dim, w.vV.as<double>(), c.d_Xs.as<double>(), w.Gv.as<double>(),
                                     nullptr, s);
                    launch_var_self_grad(nc, c.dim, w.X.as<double>(), d2, c.zeta, w.Gv.as<double>(), s);
                    launch_soap_bwd(sp, nc, centres, w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                                    w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.Gv.as<double>(),
                                    w.norm.as<double>(), w.cbuf.as<double>(), w.pairFv.as<double>(), s);
                }
            }
            if (grad) {
                if (f64fused) {
                    // G is complete
                } else if (m->precision == GAPX_PREC_FP64) {
                    launch_gemm2_f64(tu.g2, tu.parts, nc, c.M, c.dim, w.W.as<double>(), c.d_Xs.as<double>(),
                                     w.G.as<double>(), w.Gparts.as<double>(), s);
                } else if (fused[k]) {
                    // G is complete (summed with the energy rows)
                } else {
                    launch_gemm2_split(m->precision, nc, c.M, c.dim, c.dpad, c.Mpad, tu.g2lds, w.Wf.as<float>(),
                                       c.d_XsT_hi.p,
                                       c.d_XsT_lo.p, c.wscale, w.G.as<double>(), s);
                }
                const int t3 = marks ? tm.mark(s) : -1;
                launch_soap_bwd(sp, nc, centres, w.Z.as<int>(), w.nb_off.as<int>(), w.nb_j.as<int>(),
                                w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.G.as<double>(),
                                w.norm.as<double>(), w.cbuf.as<double>(), w.pairF.as<double>(), s);
                if (marks) { tm.stage(3, t2, t3); tm.stage(4, t3, tm.mark(s)); }
            }
        }
        const int tg = marks ? tm.mark(s) : -1;
        launch_gather(N, w.Z.as<int>(), m->d_e0.as<double>(), m->E_scale, w.nb_off.as<int>(), w.nb_rev.as<int>(),
                      w.nb_d.as<double>(), w.pairF.as<double>(), pairE, w.esite.as<double>(), o_f, o_le,
                      o_lv, cmask, S == 1 ? w.spart.as<double>() : nullptr, s);
        if (S == 1) launch_struct_finish(N, w.spart.as<double>(), o_e, o_v, s, w.nl_flags.as<int>(), flag_slot);
        else launch_struct_reduce(S, w.atom_off.as<int>(), o_le, o_lv, o_e, o_v, s, w.nl_flags.as<int>(), flag_slot);
        if (want_var)
            launch_gather_var(N, w.nb_off.as<int>(), w.nb_rev.as<int>(), w.pairFv.as<double>(),
                              has_d2b_var ? w.pairVE.as<double>() : nullptr, w.varsite.as<double>(), opt->d_gvg,
                              opt->d_lgv, s);
        if (marks) tm.stage(5, tg, tm.mark(s));
    };
    eg.var_valid = want_var;
    eg.var_P = P;

    if (use_graph) {
        // replay the captured sequence when nothing it depends on has changed
        std::vector<long long> key = {S, N, P, m->precision, (long long)want, (long long)(uintptr_t)o_e,
                                      (long long)(uintptr_t)o_v, (long long)(uintptr_t)o_f, (long long)(uintptr_t)o_le,
                                      (long long)(uintptr_t)o_lv, alloc_generation(), spec ? 1 : 0};
        if (spec)
            key.insert(key.end(), {(long long)(uintptr_t)d_pos, pl.NC, pl.cap1, pl.cap2, pl.capT1, pl.dev_geo ? 1 : 0});
        long long es;
        memcpy(&es, &m->E_scale, 8);
        key.push_back(es);
        for (size_t k = 0; k < NCOORD; k++)
            key.insert(key.end(), {(long long)eg.h_centres[k].size(), tunes[k].g1, tunes[k].g2, tunes[k].parts,
                                   tunes[k].fparts, tunes[k].g1lds, tunes[k].g2lds, tunes[k].f64parts});
        if (!eg.gexec || key != eg.gkey) {
            if (eg.gexec) {
                hipGraphExecDestroy(eg.gexec);
                eg.gexec = nullptr;
            }
            eg.gkey.clear();
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_all(false);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(s, &g);
            if (ce != hipSuccess || !g) return fail(GAPX_E_DEVICE, "graph capture: %s", hipGetErrorString(ce));
            const hipError_t ie = hipGraphInstantiate(&eg.gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ie != hipSuccess) {
                eg.gexec = nullptr;
                return fail(GAPX_E_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
            }
            eg.gkey = key;
        }
        HIPCHK(hipGraphLaunch(eg.gexec, s));
    } else {
        launch_all(tm.on);
    }
    HIPCHK(hipGetLastError());
    if (spec) {   // the flags (and this call's extents) are read when the caller synchronises
        const size_t need = 16 + (pl.dev_geo ? 48 * (size_t)S : 0);
        Engine::Pending pc;
        pc.S = S;
        pc.off.assign(atom_off, atom_off + S + 1);
        pc.Z.assign(Z, Z + N);
        pc.lats.assign(lats, lats + 9 * (size_t)S);
        pc.pbcs.assign(pbcs, pbcs + 3 * (size_t)S);
        pc.d_pos = d_pos;
        pc.nl_cutoff = nl_cutoff;
        pc.want = want;
        pc.d_e = d_e; pc.d_le = d_le; pc.d_f = d_f; pc.d_v = d_v; pc.d_lv = d_lv;
        pc.slot = eg.checks_used;
        pc.extents = pl.dev_geo;
        char *slot = eg.checks.as<char>() + pc.slot;
        if (!flag_slot) HIPCHK(hipMemcpyAsync(slot, w.nl_flags.p, 16, hipMemcpyDeviceToHost, s));
        if (pl.dev_geo) HIPCHK(hipMemcpyAsync(slot + 16, w.minmax.p, 48 * (size_t)S, hipMemcpyDeviceToHost, s));
        eg.checks_used += (need + 15) & ~(size_t)15;
        eg.pending.push_back(std::move(pc));
    }
    if (tm.on) tm.end();
    return GAPX_OK;
}

// Read the pending calls' flags (after one synchronisation), refresh the capacities and
// extents from them, and redo in order every call from the first that overflowed.
static int resolve_pending_list(gapx_model *m, Engine &eg, std::vector<Engine::Pending> *) {
    if (eg.pending.empty()) return GAPX_OK;
    HIPCHK(hipStreamSynchronize(eg.stream));
    eg.staging_dirty = false;
    std::vector<Engine::Pending> todo;
    todo.swap(eg.pending);
    int first_bad = -1;
    for (size_t k = 0; k < todo.size(); k++) {
        const char *slot = eg.checks.as<char>() + todo[k].slot;
        int fl[4];
        memcpy(fl, slot, 16);
        update_nl_caps(eg, fl);
        if (todo[k].extents) {
            eg.h_mm.resize(6 * (size_t)todo[k].S);
            memcpy(eg.h_mm.data(), slot + 16, 48 * (size_t)todo[k].S);
        }
        if (fl[0] != 0 && first_bad < 0) first_bad = (int)k;
    }
    eg.checks_used = 0;
    if (first_bad < 0) return GAPX_OK;
    eg.force_exact = true;
    int st = GAPX_OK;
    for (size_t k = first_bad; k < todo.size() && st == GAPX_OK; k++) {
        const Engine::Pending &p = todo[k];
        st = eval_batch(m, eg, p.S, p.off.data(), p.d_pos, nullptr, p.Z.data(), p.lats.data(), p.pbcs.data(),
                        p.nl_cutoff, p.want, p.d_e, p.d_le, p.d_f, p.d_v, p.d_lv);
    }
    eg.force_exact = false;
    CHK(st);
    HIPCHK(hipStreamSynchronize(eg.stream));
    return GAPX_OK;
}

int gapx::resolve_pending(gapx_model *m, Engine &eg) { return resolve_pending_list(m, eg, nullptr); }

extern "C" int gapx_calc_device(gapx_model *m, int N, const double *d_pos, const int *Z, const double *lat,
                                const int *pbc, double nl_cutoff, unsigned want, double *d_e, double *d_le,
                                double *d_f, double *d_v, double *d_lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    const int off[2] = {0, N};
    return eval_batch(m, *m->eng, 1, off, d_pos, nullptr, Z, lat, pbc, nl_cutoff, want, d_e, d_le, d_f, d_v, d_lv);
}

extern "C" int gapx_calc_batch_device(gapx_model *m, int n_struct, const int *atom_off, const double *d_pos,
                                      const int *Z, const double *lattices, const int *pbcs, double nl_cutoff,
                                      unsigned want, double *d_e, double *d_le, double *d_f, double *d_v,
                                      double *d_lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    return eval_batch(m, *m->eng, n_struct, atom_off, d_pos, nullptr, Z, lattices, pbcs, nl_cutoff, want, d_e, d_le,
                      d_f, d_v, d_lv);
}

extern "C" int gapx_sync(gapx_model *m) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    HIPCHK(hipSetDevice(m->device));
    CHK(resolve_pending(m, *m->eng));
    HIPCHK(hipStreamSynchronize(m->eng->stream));
    return GAPX_OK;
}

// Host-buffer evaluation of S structures on engine eg: positions go through
// pinned staging, results come back into the caller's arrays.
int gapx::calc_host(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *pos, const int *Z,
                     const double *lats, const int *pbcs, double nl_cutoff, unsigned want, double *e, double *le,
                     double *f, double *v, double *lv, const ExtNL *ext, const CalcOpts *opt, double *h_epc,
                     double *h_lgv, double *h_gvg) {
    if ((want & GAPX_WANT_E) && !e) return fail(GAPX_E_ARG, "energy requested but e is NULL");
    if ((want & GAPX_WANT_F) && !f) return fail(GAPX_E_ARG, "forces requested but f is NULL");
    if ((want & GAPX_WANT_V) && !v) return fail(GAPX_E_ARG, "virial requested but virial is NULL");
    if ((want & GAPX_WANT_LOCAL_E) && !le) return fail(GAPX_E_ARG, "local_e requested but NULL");
    if ((want & GAPX_WANT_LOCAL_V) && !lv) return fail(GAPX_E_ARG, "local_virial requested but NULL");
    CHK(check_batch(S, atom_off, Z, lats, pbcs));
    const int N = atom_off[S];
    if (N > 0 && !pos) return fail(GAPX_E_ARG, "positions are NULL");
    HIPCHK(hipSetDevice(m->device));
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    CHK(w.pos.ensure((size_t)N * 24 + 24));
    CHK(ensure_outputs(w, S, N));   // before taking the output pointers below
    // one pinned block: positions in, then results out
    const size_t bpos = (size_t)N * 24, be = (size_t)S * 8, bv = (size_t)S * 72, bf = (size_t)N * 24,
                 ble = (size_t)N * 8, blv = (size_t)N * 72;
    // option outputs: device scratch read back with the rest (sized before any copy is queued)
    CalcOpts o2;
    const size_t bepc = m->coords.size() * 8, blgv = (size_t)N * 8, bgvg = (size_t)N * 24;
    size_t bopt = 0;
    if (opt) {
        o2 = *opt;
        o2.d_epc = o2.d_lgv = o2.d_gvg = nullptr;
        if (h_epc) { CHK(w.o_epc.ensure(bepc + 8)); o2.d_epc = w.o_epc.as<double>(); bopt += bepc; }
        if (opt->variance && h_lgv) { CHK(w.o_lgv.ensure(blgv + 8)); o2.d_lgv = w.o_lgv.as<double>(); bopt += blgv; }
        if (opt->variance && h_gvg) { CHK(w.o_gvg.ensure(bgvg + 8)); o2.d_gvg = w.o_gvg.as<double>(); bopt += bgvg; }
    }
    CHK(eg.stage_io.ensure(std::max(bpos, be + bv + bf + ble + blv + bopt) + 64));
    char *h = eg.stage_io.as<char>();
    if (N) {
        memcpy(h, pos, bpos);
        HIPCHK(hipMemcpyAsync(w.pos.p, h, bpos, hipMemcpyHostToDevice, s));
    }
    CHK(eval_batch(m, eg, S, atom_off, w.pos.as<double>(), pos, Z, lats, pbcs, nl_cutoff,
                   want | GAPX_WANT_E | GAPX_WANT_V | GAPX_WANT_F | GAPX_WANT_LOCAL_E | GAPX_WANT_LOCAL_V,
                   w.e.as<double>(), w.le.as<double>(), w.f.as<double>(), w.vir.as<double>(), w.lv.as<double>(),
                   ext, opt ? &o2 : nullptr));
    CHK(resolve_pending(m, eg));   // a speculative call that overflowed is redone before its results are read
    char *o = h;
    auto get = [&](bool on, const DBuf &src, size_t bytes, char **dst) -> int {
        *dst = nullptr;
        if (!on || !bytes) return GAPX_OK;
        HIPCHK(hipMemcpyAsync(o, src.p, bytes, hipMemcpyDeviceToHost, s));
        *dst = o;
        o += bytes;
        return GAPX_OK;
    };
    char *pe, *pv, *pf, *ple, *plv;
    CHK(get(want & GAPX_WANT_E, w.e, be, &pe));
    CHK(get(want & GAPX_WANT_V, w.vir, bv, &pv));
    CHK(get(want & GAPX_WANT_F, w.f, bf, &pf));
    CHK(get(want & GAPX_WANT_LOCAL_E, w.le, ble, &ple));
    CHK(get(want & GAPX_WANT_LOCAL_V, w.lv, blv, &plv));
    char *pepc = nullptr, *plgv = nullptr, *pgvg = nullptr;
    if (opt) {
        CHK(get(o2.d_epc != nullptr, w.o_epc, bepc, &pepc));
        CHK(get(o2.d_lgv != nullptr, w.o_lgv, blgv, &plgv));
        CHK(get(o2.d_gvg != nullptr, w.o_gvg, bgvg, &pgvg));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (pepc) memcpy(h_epc, pepc, bepc);
    if (plgv) memcpy(h_lgv, plgv, blgv);
    if (pgvg) memcpy(h_gvg, pgvg, bgvg);
    if (pe) memcpy(e, pe, be);
    if (pv) memcpy(v, pv, bv);
    if (pf) memcpy(f, pf, bf);
    if (ple) memcpy(le, ple, ble);
    if (plv) memcpy(lv, plv, blv);
    return GAPX_OK;
}

extern "C" int gapx_calc(gapx_model *m, int N, const double *pos, const int *Z, const double *lat, const int *pbc,
                         double nl_cutoff, unsigned want, double *e, double *le, double *f, double *v, double *lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    const int off[2] = {0, N};
    return calc_host(m, *m->eng, 1, off, pos, Z, lat, pbc, nl_cutoff, want, e, le, f, v, lv);
}

static int make_opts(gapx_model *m, unsigned want, const unsigned char *atom_mask, int only_descriptor,
                     double reg, double *epc, double *lgv, double *gvg, CalcOpts &o) {
    if (only_descriptor < 0) return fail(GAPX_E_ARG, "only_descriptor must be >= 0 (got %d)", only_descriptor);
    o.mask = atom_mask;
    o.only_coord = only_descriptor;
    o.d_epc = epc;
    o.variance = (want & GAPX_WANT_VARIANCE) || lgv || gvg;
    o.reg = reg;
    o.d_lgv = lgv;
    o.d_gvg = gvg;
    (void)m;
    return GAPX_OK;
}

extern "C" int gapx_calc_ex(gapx_model *m, int N, const double *pos, const int *Z, const double *lat, const int *pbc,
                            double nl_cutoff, unsigned want, const unsigned char *atom_mask, int only_descriptor,
                            double variance_regularisation, double *e, double *le, double *f, double *v, double *lv,
                            double *energy_per_coordinate, double *local_gap_variance,
                            double *gap_variance_gradient) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    CalcOpts o;
    CHK(make_opts(m, want, atom_mask, only_descriptor, variance_regularisation, energy_per_coordinate,
                  local_gap_variance, gap_variance_gradient, o));
    const int off[2] = {0, N};
    return calc_host(m, *m->eng, 1, off, pos, Z, lat, pbc, nl_cutoff, want, e, le, f, v, lv, nullptr, &o,
                     energy_per_coordinate, local_gap_variance, gap_variance_gradient);
}

extern "C" int gapx_calc_device_ex(gapx_model *m, int N, const double *d_pos, const int *Z, const double *lat,
                                   const int *pbc, double nl_cutoff, unsigned want, const unsigned char *atom_mask,
                                   int only_descriptor, double variance_regularisation, double *d_e, double *d_le,
                                   double *d_f, double *d_v, double *d_lv, double *d_energy_per_coordinate,
                                   double *d_local_gap_variance, double *d_gap_variance_gradient) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    CalcOpts o;
    CHK(make_opts(m, want, atom_mask, only_descriptor, variance_regularisation, d_energy_per_coordinate,
                  d_local_gap_variance, d_gap_variance_gradient, o));
    const int off[2] = {0, N};
    return eval_batch(m, *m->eng, 1, off, d_pos, nullptr, Z, lat, pbc, nl_cutoff, want, d_e, d_le, d_f, d_v, d_lv,
                      nullptr, &o);
}

// The per-instance variances of coordinate `coord` from the last calc that formed them,
// in the reference's instance order (print_gap_variance, IPModel_GAP.f95:506-516):
// soap, one per centre atom (atoms[2k] = centre, atoms[2k+1] = -1, cutoff_sq = 1);
// distance_2b, one per directed pair (i, j) with i ascending and j in i's neighbour order.
extern "C" int gapx_descriptor_variances(gapx_model *m, int coord, int max_entries, double *variance,
                                         double *cutoff_sq, int *atoms, int *n_entries) {
    if (!m || !n_entries) return fail(GAPX_E_ARG, "NULL argument");
    if (coord < 0 || coord >= (int)m->coords.size()) return fail(GAPX_E_ARG, "coordinate %d out of range", coord);
    Engine &eg = *m->eng;
    if (!eg.var_valid) return fail(GAPX_E_ARG, "no variance formed by the last calc (ask for GAPX_WANT_VARIANCE)");
    HIPCHK(hipSetDevice(m->device));
    hipStream_t s = eg.stream;
    HIPCHK(hipStreamSynchronize(s));
    const Coord &c = *m->coords[coord];
    Workspace &w = eg.ws;
    int k = 0;
    auto put = [&](double var, double c2, int a, int b) {
        if (k < max_entries) {
            if (variance) variance[k] = var;
            if (cutoff_sq) cutoff_sq[k] = c2;
            if (atoms) { atoms[2 * k] = a; atoms[2 * k + 1] = b; }
        }
        k++;
    };
    if (!w.var_inst[coord].p) {   // a coordinate the last calc skipped (only_descriptor)
        *n_entries = 0;
        return GAPX_OK;
    }
    if (c.kind == 1) {
        const auto &cen = eg.h_centres[coord];
        std::vector<double> v(cen.size());
        if (!cen.empty()) HIPCHK(hipMemcpy(v.data(), w.var_inst[coord].p, cen.size() * 8, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < cen.size(); q++) put(v[q], 1.0, cen[q], -1);
    } else {
        const int N = (int)eg.h_Z.size(), P = eg.var_P;
        std::vector<int> off(N + 1), nj(P + 1);
        std::vector<double> v(P + 1), c2(P + 1);
        HIPCHK(hipMemcpy(off.data(), w.nb_off.p, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost));
        if (P) {
            HIPCHK(hipMemcpy(nj.data(), w.nb_j.p, (size_t)P * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(v.data(), w.var_inst[coord].p, (size_t)P * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(c2.data(), w.cut2_inst[coord].p, (size_t)P * 8, hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < N; i++)
            for (int p = off[i]; p < off[i + 1]; p++)
                if (!std::isnan(v[p])) put(v[p], c2[p], i, nj[p]);
    }
    *n_entries = k;
    return GAPX_OK;
}

// LAMMPS pair style entry (quip_lammps_wrapper.f95:30-156).  The connection is
// filled as the reference fills at%connect (:104-144): for each listed local
// atom i and each LAMMPS neighbour j (1-based in quip_neigh) with i <= j, j is
// appended to neighbour1(i) with zero shift and r = |x_i - x_j|, and (i, that
// index) to neighbour2(j); descriptors are centred on the ilist atoms only
// (atom_mask_name=local, :147), and e0 is added for those atoms only
// (IPModel_GAP.f95:577-589).
extern "C" int gapx_calc_lammps(gapx_model *m, int nlocal, int nghost, const int *Zs, const int *lmptag, int inum,
                                int sum_num_neigh, const int *ilist, const int *num_neigh, const int *neigh,
                                const double *lattice, const double *x, double *e, double *le, double *vir,
                                double *lv, double *f) {
    (void)lmptag;
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (nlocal < 0 || nghost < 0 || inum < 0 || sum_num_neigh < 0)
        return fail(GAPX_E_ARG, "negative count (nlocal %d nghost %d inum %d sum_num_neigh %d)", nlocal, nghost, inum,
                    sum_num_neigh);
    const long Nl = (long)nlocal + nghost;
    if (Nl > INT32_MAX / 16) return fail(GAPX_E_ARG, "too many atoms (%ld)", Nl);
    const int N = (int)Nl;
    if (!e || !le || !vir || !lv || !f) return fail(GAPX_E_ARG, "an output array is NULL");
    if (nlocal == 0) {   // this region is vacuum (:150-155)
        *e = 0.0;
        for (int k = 0; k < 9; k++) vir[k] = 0.0;
        for (long i = 0; i < Nl; i++) {
            le[i] = 0.0;
            for (int k = 0; k < 3; k++) f[3 * i + k] = 0.0;
            for (int k = 0; k < 9; k++) lv[9 * i + k] = 0.0;
        }
        return GAPX_OK;
    }
    if (!Zs || !x) return fail(GAPX_E_ARG, "atomic_numbers / quip_x are NULL");
    if (inum > 0 && (!ilist || !num_neigh)) return fail(GAPX_E_ARG, "ilist / quip_num_neigh are NULL");
    if (sum_num_neigh > 0 && !neigh) return fail(GAPX_E_ARG, "quip_neigh is NULL");
    ExtNL c;
    c.local.assign(N, 0);
    std::vector<int> n1c(N, 0), n2c(N, 0);
    long nn = 0;
    for (int ni = 0; ni < inum; ni++) {   // validate and count
        const int i = ilist[ni];
        if (i < 0 || i >= N) return fail(GAPX_E_ARG, "ilist[%d] = %d outside [0, %d)", ni, i, N);
        if (c.local[i]) return fail(GAPX_E_ARG, "atom %d listed twice in ilist", i);
        c.local[i] = 1;
        if (num_neigh[ni] < 0) return fail(GAPX_E_ARG, "quip_num_neigh[%d] < 0", ni);
        for (int n = 0; n < num_neigh[ni]; n++, nn++) {
            if (nn >= sum_num_neigh) return fail(GAPX_E_ARG, "quip_num_neigh adds up to more than sum_num_neigh");
            const int j = neigh[nn] - 1;
            if (j < 0 || j >= N) return fail(GAPX_E_ARG, "quip_neigh[%ld] = %d outside [1, %d]", nn, j + 1, N);
            if (i <= j) {
                n1c[i]++;
                n2c[j]++;
            }
        }
    }
    c.n1_off.assign(N + 1, 0);
    c.n2_off.assign(N + 1, 0);
    for (int i = 0; i < N; i++) {
        c.n1_off[i + 1] = c.n1_off[i] + n1c[i];
        c.n2_off[i + 1] = c.n2_off[i] + n2c[i];
    }
    const int T = c.n1_off[N];
    c.n1_j.resize(T);
    c.n1_r.resize(T);
    c.n1_rev.resize(T);
    c.n2_i.resize(T);
    c.n2_k.resize(T);
    std::fill(n1c.begin(), n1c.end(), 0);
    std::fill(n2c.begin(), n2c.end(), 0);
    nn = 0;
    for (int ni = 0; ni < inum; ni++) {   // fill in the reference's append order
        const int i = ilist[ni];
        for (int n = 0; n < num_neigh[ni]; n++, nn++) {
            const int j = neigh[nn] - 1;
            if (i > j) continue;
            const double dx = x[3 * (size_t)i] - x[3 * (size_t)j], dy = x[3 * (size_t)i + 1] - x[3 * (size_t)j + 1],
                         dz = x[3 * (size_t)i + 2] - x[3 * (size_t)j + 2];
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r < 1.0e-10) return fail(GAPX_E_GEOMETRY, "atoms %d and %d overlap exactly", i + 1, j + 1);
            const int k1 = n1c[i]++, e1 = c.n1_off[i] + k1;
            c.n1_j[e1] = j;
            c.n1_r[e1] = r;
            const int k2 = n2c[j]++;
            c.n2_i[c.n2_off[j] + k2] = i;
            c.n2_k[c.n2_off[j] + k2] = k1;
            c.n1_rev[e1] = k2;
        }
    }
    static const double eye[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const int off[2] = {0, N}, pbc[3] = {0, 0, 0};
    const unsigned all = GAPX_WANT_E | GAPX_WANT_F | GAPX_WANT_V | GAPX_WANT_LOCAL_E | GAPX_WANT_LOCAL_V;
    return calc_host(m, *m->eng, 1, off, x, Zs, lattice ? lattice : eye, pbc, 0.0, all, e, le, f, vir, lv, &c);
}

extern "C" int gapx_calc_batch(gapx_model *m, int n_struct, const int *atom_off, const double *pos, const int *Z,
                               const double *lattices, const int *pbcs, double nl_cutoff, unsigned want,
                               double *energies, double *local_e, double *f, double *virials, double *local_virial) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    return calc_host(m, *m->eng, n_struct, atom_off, pos, Z, lattices, pbcs, nl_cutoff, want, energies, local_e, f,
                     virials, local_virial);
}

// ---------------------------------------------------------------------------
// neighbour list only
// ---------------------------------------------------------------------------
extern "C" int gapx_connect(int device, int N, const double *pos, const double *lat, const int *pbc, double cutoff,
                            gapx_nl **out) {
    if (!out || (N > 0 && !pos) || !lat || !pbc || N < 0) return fail(GAPX_E_ARG, "bad arguments");
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    gapx_model tmp;   // no coordinates: only the neighbour-list part of an engine is used
    tmp.device = device;
    CHK(new_engine(&tmp, tmp.eng));
    Engine &eg = *tmp.eng;
    Workspace &w = eg.ws;
    const int off[2] = {0, N};
    std::vector<int> Z0((size_t)N, 0);
    CHK(check_batch(1, off, Z0.data(), lat, pbc));
    CHK(w.pos.ensure((size_t)N * 24 + 24));
    if (N > 0) HIPCHK(hipMemcpy(w.pos.p, pos, (size_t)N * 24, hipMemcpyHostToDevice));
    CHK(stage_inputs(&tmp, eg, 1, off, Z0.data()));
    int P = 0;
    CHK(build_nl(&tmp, eg, 1, off, w.pos.as<double>(), pos, lat, pbc, cutoff, &P));
    HIPCHK(hipStreamSynchronize(eg.stream));
    std::unique_ptr<gapx_nl> nl(new gapx_nl());
    nl->N = N;
    const int T1 = eg.n1_total_last;
    nl->n1_off.resize(N + 1); nl->n2_off.resize(N + 1); nl->map_shift.resize(3 * (size_t)N);
    nl->n1_j.resize(T1); nl->n1_shift.resize(3 * (size_t)T1); nl->n1_r.resize(T1);
    HIPCHK(hipMemcpy(nl->n1_off.data(), w.n1_off.p, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nl->n2_off.data(), w.n2_off.p, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost));
    if (N) HIPCHK(hipMemcpy(nl->map_shift.data(), w.map_shift.p, (size_t)N * 12, hipMemcpyDeviceToHost));
    if (T1) {
        HIPCHK(hipMemcpy(nl->n1_j.data(), w.n1_j.p, (size_t)T1 * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nl->n1_shift.data(), w.n1_shift.p, (size_t)T1 * 12, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nl->n1_r.data(), w.n1_r.p, (size_t)T1 * 8, hipMemcpyDeviceToHost));
    }
    const int T2 = nl->n2_off[N];
    nl->n2_j.resize(T2); nl->n2_idx.resize(T2);
    if (T2) {
        HIPCHK(hipMemcpy(nl->n2_j.data(), w.n2_is.p, (size_t)T2 * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nl->n2_idx.data(), w.n2_ks.p, (size_t)T2 * 4, hipMemcpyDeviceToHost));
    }
    for (int a = 0; a < 3; a++) nl->cells[a] = eg.h_geo[0].cells[a];
    HIPCHK(hipGetLastError());
    *out = nl.release();
    return GAPX_OK;
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
    Engine &eg = *m->eng;
    CHK(resolve_pending(m, eg));
    Workspace &w = eg.ws;
    hipStream_t s = eg.stream;
    const int off[2] = {0, N};
    CHK(check_batch(1, off, Z, lat, pbc));
    if (N > 0 && !pos) return fail(GAPX_E_ARG, "positions are NULL");
    CHK(w.pos.ensure((size_t)N * 24 + 24));
    CHK(eg.stage_io.ensure((size_t)N * 24 + 64));
    if (N) {
        memcpy(eg.stage_io.p, pos, (size_t)N * 24);
        HIPCHK(hipMemcpyAsync(w.pos.p, eg.stage_io.p, (size_t)N * 24, hipMemcpyHostToDevice, s));
    }
    CHK(stage_inputs(m, eg, 1, off, Z));
    int P = 0;
    CHK(build_nl(m, eg, 1, off, w.pos.as<double>(), pos, lat, pbc, nl_cutoff > 0 ? nl_cutoff : c.rc, &P));
    const int nc = (int)eg.h_centres[coord].size();
    *n_desc = nc;
    if (nc == 0) return GAPX_OK;
    const int KLM = c.S * c.n * (c.L + 1) * (c.L + 1);
    CHK(w.X.ensure((size_t)nc * c.dim * 8));
    CHK(w.norm.ensure((size_t)nc * 8));
    CHK(w.cbuf.ensure((size_t)nc * KLM * 8));
    launch_soap_fwd(soap_dev(c), nc, w.centres[coord].as<int>(), w.Z.as<int>(), w.nb_off.as<int>(),
                    w.nb_j.as<int>(), w.nb_r.as<double>(), w.nb_d.as<double>(), w.X.as<double>(), w.norm.as<double>(),
                    w.cbuf.as<double>(), s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(X, w.X.p, (size_t)nc * c.dim * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(desc_atom, eg.h_centres[coord].data(), (size_t)nc * 4);
    return GAPX_OK;
}

// mixed-precision sparse-point copies are built lazily the first time a model
// runs in a split precision
int gapx::ensure_split(gapx_model *m, Coord &c, int prec, hipStream_t s) {
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
    launch_split_matrix(prec, c.M, c.dim, c.dpad, c.d_Xs.as<double>(), c.dim, false, c.d_Xs_hi.p, c.d_Xs_lo.p, s);
    launch_split_matrix(prec, c.M, c.dim, c.Mpad, c.d_Xs.as<double>(), c.dim, true, c.d_XsT_hi.p, c.d_XsT_lo.p, s);
    HIPCHK(hipGetLastError());
    c.split_prec = prec;
    return GAPX_OK;
}
