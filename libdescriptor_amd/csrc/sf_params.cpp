// Parameter sets of the symmetry-function descriptor: the vector constructor
// (src/SymmetryFunctions.cpp:429-450 via src/Descriptors.cpp:579-590), the
// parameter-file reader (src/SymmetryFunctions.cpp:262-407 with the tokenizer
// of include/file_io_utils.hpp:53-186), and the slot layout
// (SymmetryFunctions::add_descriptor, src/SymmetryFunctions.cpp:43-65).
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ldsf.h"
#include "ldsf_internal.h"

namespace {

thread_local std::string g_last_error;

int kind_code(const std::string &name)
{
    if (name == "g1") return ldsf::G1;
    if (name == "g2") return ldsf::G2;
    if (name == "g3") return ldsf::G3;
    if (name == "g4") return ldsf::G4;
    if (name == "g5") return ldsf::G5;
    return 0;
}

// ---- tokenizer with the semantics of include/file_io_utils.hpp ----

// next line that is neither empty nor a comment; false at end of file (the
// reference loops forever there, file_io_utils.hpp:53-59)
bool next_data_line(std::ifstream &f, std::string &line)
{
    line.clear();
    while (line.empty() || line[0] == '#') {
        if (!std::getline(f, line)) return false;
    }
    return true;
}

// parse_int_params: tokens that std::stoi rejects are skipped (:70-96)
bool parse_ints(const std::string &line, std::vector<int> &out, int n)
{
    std::stringstream ss(line);
    std::string tok;
    int got = 0;
    while (got < n && ss >> tok) {
        try {
            out.push_back(std::stoi(tok));
            ++got;
        } catch (std::invalid_argument &) {
        } catch (std::out_of_range &) {
            return false;
        }
    }
    return got == n;
}

// parse_double_params (:108-134)
bool parse_doubles(const std::string &line, std::vector<double> &out, int n)
{
    std::stringstream ss(line);
    std::string tok;
    int got = 0;
    while (got < n && ss >> tok) {
        try {
            out.push_back(std::stod(tok));
            ++got;
        } catch (std::invalid_argument &) {
        } catch (std::out_of_range &) {
            return false;
        }
    }
    return got == n;
}

// parse_string_params (:145-160)
bool parse_strings(const std::string &line, std::vector<std::string> &out, int n)
{
    std::stringstream ss(line);
    std::string tok;
    int got = 0;
    while (got < n && ss >> tok) {
        out.push_back(tok);
        ++got;
    }
    return got == n;
}

// parse_bool_params (:171-186)
bool parse_bools(const std::string &line, std::vector<bool> &out, int n)
{
    std::stringstream ss(line);
    std::string tok;
    int got = 0;
    while (got < n && ss >> tok) {
        if (tok == "true" || tok == "True" || tok == "TRUE") { out.push_back(true); ++got; }
        else if (tok == "false" || tok == "False" || tok == "FALSE") { out.push_back(false); ++got; }
    }
    return got == n;
}

}  // namespace

ldsf_sf::~ldsf_sf()
{
    for (auto &d : dev)
        if (d.mem && d.release) d.release(d.device, d.mem);
}

namespace ldsf {

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

namespace {
thread_local long g_rows_limit = -1;
}

long host_rows_limit() { return g_rows_limit; }

HostRowsLimit::HostRowsLimit(long rows) : saved(g_rows_limit) { g_rows_limit = rows; }

HostRowsLimit::~HostRowsLimit() { g_rows_limit = saved; }

int finalize_params(ldsf_sf *sf)
{
    const int ns = (int)sf->species.size();
    const int npairs = ns * (ns + 1) / 2;
    sf->start.clear();
    sf->k2.clear(); sf->slot2.clear(); sf->str2.clear(); sf->a2.clear(); sf->b2.clear();
    sf->k3.clear(); sf->slot3.clear(); sf->str3.clear(); sf->zeta3.clear(); sf->lam3.clear();
    sf->eta3.clear(); sf->coef3.clear(); sf->izeta3.clear();
    sf->pidx2.clear(); sf->pidx3.clear();
    sf->has_three_body = 0;
    int w = 0, np = 0;
    for (size_t p = 0; p < sf->kind.size(); ++p) {
        int k = sf->kind[p], R = sf->rows[p], C = sf->cols[p];
        if (R < 0 || C < 0) return fail(LDSF_ERR_ARG, "negative parameter block size");
        const std::vector<double> &P = sf->par[p];
        sf->start.push_back(w);
        if (two_body(k)) {
            if ((k == G2 && C < 2) || (k == G3 && C < 1))
                return fail(LDSF_ERR_ARG, "g2 needs 2 parameter columns (eta, Rs), g3 needs 1 (kappa)");
            for (int q = 0; q < R; ++q) {
                sf->k2.push_back(k);
                sf->a2.push_back(C > 0 ? P[q * C + 0] : 0.0);
                sf->b2.push_back(C > 1 ? P[q * C + 1] : 0.0);
                sf->slot2.push_back(w + q);
                sf->str2.push_back(sf->species_resolved ? R : 0);
                sf->pidx2.push_back(np + q * C);
            }
            w += R * (sf->species_resolved ? ns : 1);
        } else if (three_body(k)) {
            if (C < 3) return fail(LDSF_ERR_ARG, "g4/g5 need 3 parameter columns (zeta, lambda, eta)");
            sf->has_three_body = 1;
            for (int q = 0; q < R; ++q) {
                double zeta = P[q * C + 0];
                sf->k3.push_back(k);
                sf->zeta3.push_back(zeta);
                sf->lam3.push_back(P[q * C + 1]);
                sf->eta3.push_back(P[q * C + 2]);
                sf->coef3.push_back(std::pow(2, 1 - zeta));
                int iz = -1;
                if (zeta >= 0 && zeta <= kMaxIntZeta && zeta == std::floor(zeta)) iz = (int)zeta;
                sf->izeta3.push_back(iz);
                sf->slot3.push_back(w + q);
                sf->str3.push_back(sf->species_resolved ? R : 0);
                sf->pidx3.push_back(np + q * C);
            }
            w += R * (sf->species_resolved ? npairs : 1);
        } else {
            return fail(LDSF_ERR_ARG, "unknown symmetry function type");
        }
        np += R * C;
    }
    sf->width = w;
    sf->n_params = np;
    // the device table's host image
    PackedTable &T = sf->packed;
    T.d.clear();
    T.i.clear();
    auto put_d = [&](const std::vector<double> &v) {
        const size_t o = T.d.size();
        T.d.insert(T.d.end(), v.begin(), v.end());
        return o;
    };
    auto put_i = [&](const std::vector<int> &v) {
        const size_t o = T.i.size();
        T.i.insert(T.i.end(), v.begin(), v.end());
        return o;
    };
    T.o_rc = put_d(sf->rcut);
    T.o_a2 = put_d(sf->a2);
    T.o_b2 = put_d(sf->b2);
    T.o_z3 = put_d(sf->zeta3);
    T.o_l3 = put_d(sf->lam3);
    T.o_e3 = put_d(sf->eta3);
    T.o_c3 = put_d(sf->coef3);
    T.o_k2 = put_i(sf->k2);
    T.o_s2 = put_i(sf->slot2);
    T.o_t2 = put_i(sf->str2);
    T.o_k3 = put_i(sf->k3);
    T.o_iz = put_i(sf->izeta3);
    T.o_s3 = put_i(sf->slot3);
    T.o_t3 = put_i(sf->str3);
    T.o_p2 = put_i(sf->pidx2);
    T.o_p3 = put_i(sf->pidx3);
    return LDSF_OK;
}

}  // namespace ldsf

extern "C" {

LDSF_API const char *ldsf_last_error(void) { return g_last_error.c_str(); }

LDSF_API int ldsf_sf_create(ldsf_sf **out, int n_species, const char *const *species_names,
                            const double *cutoff_matrix, int n_types, const char *const *type_names,
                            const int *sizes, const double *params, int species_resolved)
{
    if (!out) return ldsf::fail(LDSF_ERR_ARG, "sf_create: null output");
    *out = nullptr;
    if (n_species <= 0 || !cutoff_matrix || n_types < 0 || (n_types > 0 && (!type_names || !sizes)))
        return ldsf::fail(LDSF_ERR_ARG, "sf_create: bad arguments");
    auto *sf = new ldsf_sf;
    for (int s = 0; s < n_species; ++s)
        sf->species.emplace_back(species_names && species_names[s] ? species_names[s]
                                                                   : std::to_string(s));
    sf->rcut.assign(cutoff_matrix, cutoff_matrix + n_species * n_species);
    sf->species_resolved = species_resolved ? 1 : 0;
    long off = 0;
    for (int t = 0; t < n_types; ++t) {
        int k = kind_code(type_names[t] ? type_names[t] : "");
        if (!k) {
            delete sf;
            return ldsf::fail(LDSF_ERR_ARG, std::string("unknown symmetry function type: ") +
                                                (type_names[t] ? type_names[t] : "(null)"));
        }
        int R = sizes[2 * t], C = sizes[2 * t + 1];
        sf->kind.push_back(k);
        sf->rows.push_back(R);
        sf->cols.push_back(C);
        std::vector<double> P;
        if (R > 0 && C > 0) {
            if (!params) { delete sf; return ldsf::fail(LDSF_ERR_ARG, "sf_create: null params"); }
            P.assign(params + off, params + off + (long)R * C);
        }
        sf->par.push_back(std::move(P));
        off += (long)R * C;
    }
    int rc = ldsf::finalize_params(sf);
    if (rc) { delete sf; return rc; }
    *out = sf;
    return LDSF_OK;
}

// initFromFile (src/SymmetryFunctions.cpp:266-407), behaviour kept:
//  * num_species pair lines are read ("A B cutoff"); species are the sorted
//    set of names seen; pairs never listed get cutoff 0.0 (std::map default);
//  * the per-type parameter vector is never cleared between types, so type i
//    is initialised from the first rows*cols values read so far (:335-356);
//  * normalize and descriptor_size are read and not used.
// A truncated file is an error here (the reference's reader loops forever).
LDSF_API int ldsf_sf_create_from_file(ldsf_sf **out, const char *path)
{
    if (!out || !path) return ldsf::fail(LDSF_ERR_ARG, "sf_create_from_file: null argument");
    *out = nullptr;
    std::ifstream f(path);
    if (!f.is_open()) return ldsf::fail(LDSF_ERR_FILE, std::string("Could not open file: ") + path);
    std::string line;
    std::vector<std::string> sp;
    std::vector<int> ip;
    std::vector<double> dp;
    std::vector<bool> bp;
    auto truncated = [&]() { return ldsf::fail(LDSF_ERR_FILE, std::string("unexpected end of file: ") + path); };

    if (!next_data_line(f, line)) return truncated();
    if (!parse_strings(line, sp, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all string parameters");
    sp.clear();
    if (!next_data_line(f, line)) return truncated();
    if (!parse_ints(line, ip, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all int parameters");
    int num_species = ip[0];
    ip.clear();
    if (num_species <= 0) return ldsf::fail(LDSF_ERR_FILE, "number of species must be positive");
    std::set<std::string> names;
    std::map<std::pair<std::string, std::string>, double> pair_cut;
    for (int s = 0; s < num_species; ++s) {
        if (!next_data_line(f, line)) return truncated();
        if (!parse_strings(line, sp, 2)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all string parameters");
        if (!parse_doubles(line, dp, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all double parameters");
        names.insert(sp[0]);
        names.insert(sp[1]);
        pair_cut[{sp[0], sp[1]}] = dp[0];
        pair_cut[{sp[1], sp[0]}] = dp[0];
        sp.clear();
        dp.clear();
    }
    std::vector<std::string> species(names.begin(), names.end());
    if (!next_data_line(f, line)) return truncated();
    if (!parse_ints(line, ip, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all int parameters");
    int n_types = ip[0];
    ip.clear();
    std::vector<std::string> types;
    std::map<std::string, std::pair<int, int>> type_size;
    std::map<std::string, std::vector<double>> type_par;
    std::vector<double> running;  // never cleared, as in the reference
    for (int t = 0; t < n_types; ++t) {
        if (!next_data_line(f, line)) return truncated();
        if (!parse_strings(line, sp, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all string parameters");
        types.push_back(sp[0]);
        sp.clear();
        if (!parse_ints(line, ip, 2)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all int parameters");
        type_size[types[t]] = {ip[0], ip[1]};
        ip.clear();
        for (int r = 0; r < type_size[types[t]].first; ++r) {
            if (!next_data_line(f, line)) return truncated();
            if (!parse_doubles(line, dp, type_size[types[t]].second))
                return ldsf::fail(LDSF_ERR_FILE, "Could not read all double parameters");
            for (double v : dp) running.push_back(v);
            type_par[types[t]] = running;
            dp.clear();
        }
    }
    if (!next_data_line(f, line)) return truncated();
    if (!parse_bools(line, bp, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all bool parameters");
    if (!next_data_line(f, line)) return truncated();
    if (!parse_ints(line, ip, 1)) return ldsf::fail(LDSF_ERR_FILE, "Could not read all int parameters");

    auto *sf = new ldsf_sf;
    // the matrix has num_species rows/columns even if more names were seen
    // (SymmetryFunctions.cpp:383-389); codes follow the sorted names
    if ((int)species.size() < num_species) {
        delete sf;
        return ldsf::fail(LDSF_ERR_FILE, "fewer distinct species names than declared");
    }
    species.resize(num_species);
    sf->species = species;
    sf->rcut.resize(num_species * num_species);
    for (int a = 0; a < num_species; ++a)
        for (int b = 0; b < num_species; ++b) {
            auto it = pair_cut.find({species[a], species[b]});
            sf->rcut[a * num_species + b] = it == pair_cut.end() ? 0.0 : it->second;
        }
    for (int t = 0; t < n_types; ++t) {
        int k = kind_code(types[t]);
        if (!k) { delete sf; return ldsf::fail(LDSF_ERR_FILE, "unknown symmetry function type: " + types[t]); }
        auto sz = type_size[types[t]];
        const std::vector<double> &P = type_par[types[t]];
        sf->kind.push_back(k);
        sf->rows.push_back(sz.first);
        sf->cols.push_back(sz.second);
        long need = (long)sz.first * sz.second;
        if ((long)P.size() < need) { delete sf; return ldsf::fail(LDSF_ERR_FILE, "parameter table too short"); }
        sf->par.emplace_back(P.begin(), P.begin() + need);
    }
    int rc = ldsf::finalize_params(sf);
    if (rc) { delete sf; return rc; }
    *out = sf;
    return LDSF_OK;
}

LDSF_API void ldsf_sf_destroy(ldsf_sf *sf) { delete sf; }

LDSF_API int ldsf_sf_width(const ldsf_sf *sf) { return sf ? sf->width : -1; }

LDSF_API int ldsf_sf_num_params(const ldsf_sf *sf) { return sf ? sf->n_params : -1; }

LDSF_API int ldsf_sf_num_species(const ldsf_sf *sf) { return sf ? (int)sf->species.size() : 0; }

LDSF_API const char *ldsf_sf_species_name(const ldsf_sf *sf, int k)
{
    if (!sf || k < 0 || k >= (int)sf->species.size()) return nullptr;
    return sf->species[k].c_str();
}

LDSF_API int ldsf_sf_set_precision(ldsf_sf *sf, int precision)
{
    if (!sf || (precision != LDSF_FP64 && precision != LDSF_FP32_BULK))
        return ldsf::fail(LDSF_ERR_ARG, "set_precision: bad arguments");
    sf->precision = precision;
    return LDSF_OK;
}

}  // extern "C"
