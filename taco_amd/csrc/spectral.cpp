// spectral.cpp — spectral densities J(w) given as expressions, the bath
// correlation from such a J, and the rectangle-rule Gamma (host side).
//
// The expression language is the one the reference's configuration files use
// for J (expression.hpp:14-26): the variable w, the constants pi and e, named
// parameters, + - * / ^ with unary minus, exp log sqrt abs sin cos tan
// step/heaviside of one argument and pow max min of two.  Precedence, from
// loosest: + -, then * /, then unary minus, then ^ (expression.hpp:235-253);
// ^ and unary minus group to the right.  Functions and the constants are
// matched case-insensitively, parameters by exact name.  Evaluation calls the
// platform libm, as the reference does.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "engine_internal.hpp"

using taco_engine::Error;

struct taco_spectral {
    // Program for a stack machine: operands pushed, operators applied to the
    // top of the stack.
    enum Code : unsigned char {
        PUSH, VAR_W, NEG, ADD, SUB, MUL, DIV, POW,
        F_EXP, F_LOG, F_SQRT, F_ABS, F_SIN, F_COS, F_TAN, F_STEP, F_POW, F_MAX, F_MIN
    };
    struct Ins {
        Code code;
        double value;
    };
    std::string source;
    std::vector<Ins> prog;

    double eval(double w) const {
        double st[64] = {};
        int top = 0;
        std::vector<double> big;  // deeper programs than 64 operands
        double* s = st;
        if (prog.size() > 64) {
            big.resize(prog.size());
            s = big.data();
        }
        for (const Ins& in : prog) {
            switch (in.code) {
                case PUSH: s[top++] = in.value; break;
                case VAR_W: s[top++] = w; break;
                case NEG: s[top - 1] = -s[top - 1]; break;
                case F_EXP: s[top - 1] = std::exp(s[top - 1]); break;
                case F_LOG: s[top - 1] = std::log(s[top - 1]); break;
                case F_SQRT: s[top - 1] = std::sqrt(s[top - 1]); break;
                case F_ABS: s[top - 1] = std::abs(s[top - 1]); break;
                case F_SIN: s[top - 1] = std::sin(s[top - 1]); break;
                case F_COS: s[top - 1] = std::cos(s[top - 1]); break;
                case F_TAN: s[top - 1] = std::tan(s[top - 1]); break;
                case F_STEP: s[top - 1] = s[top - 1] > 0.0 ? 1.0 : 0.0; break;
                default: {
                    const double b = s[--top], a = s[top - 1];
                    double r = 0.0;
                    switch (in.code) {
                        case ADD: r = a + b; break;
                        case SUB: r = a - b; break;
                        case MUL: r = a * b; break;
                        case DIV: r = a / b; break;
                        case POW:
                        case F_POW: r = std::pow(a, b); break;
                        case F_MAX: r = std::max(a, b); break;
                        case F_MIN: r = std::min(a, b); break;
                        default: break;
                    }
                    s[top - 1] = r;
                }
            }
        }
        return s[0];
    }
};

namespace {

using Code = taco_spectral::Code;

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

int arity_of(Code c) {
    switch (c) {
        case taco_spectral::PUSH:
        case taco_spectral::VAR_W: return 0;
        case taco_spectral::NEG:
        case taco_spectral::F_EXP:
        case taco_spectral::F_LOG:
        case taco_spectral::F_SQRT:
        case taco_spectral::F_ABS:
        case taco_spectral::F_SIN:
        case taco_spectral::F_COS:
        case taco_spectral::F_TAN:
        case taco_spectral::F_STEP: return 1;
        default: return 2;
    }
}

// Shunting-yard translation of the source into the stack program.
struct Compiler {
    struct Pending {
        enum Kind { OP, FUNC, PAREN } kind;
        Code code;  // OP: the operator; FUNC: the one-argument code (the two-argument one follows)
        int args;   // FUNC: arguments seen so far
    };
    const std::string& src;
    const std::vector<std::pair<std::string, double>>& params;
    std::vector<taco_spectral::Ins> out;
    std::vector<Pending> st;

    static int prec(Code c) {
        switch (c) {
            case taco_spectral::ADD:
            case taco_spectral::SUB: return 1;
            case taco_spectral::MUL:
            case taco_spectral::DIV: return 2;
            case taco_spectral::NEG: return 3;
            default: return 4;  // POW
        }
    }
    static bool right(Code c) { return c == taco_spectral::POW || c == taco_spectral::NEG; }

    void emit(Code c, double v = 0.0) { out.push_back({c, v}); }

    void emit_func(const Pending& f) {
        const bool two = f.code == taco_spectral::F_POW || f.code == taco_spectral::F_MAX ||
                         f.code == taco_spectral::F_MIN;
        if (f.args == 1 || f.args == 2) {
            if (two != (f.args == 2))
                throw Error(TACO_ERR_INVALID, two ? "pow/max/min take two arguments" : "function takes one argument");
            emit(f.code);
        } else {
            throw Error(TACO_ERR_INVALID, "Unsupported function arity (only 1 or 2 args supported)");
        }
    }

    void pop_to_output() {
        const Pending p = st.back();
        st.pop_back();
        if (p.kind == Pending::OP) emit(p.code);
        else if (p.kind == Pending::FUNC) emit_func(p);
    }

    static Code func_code(const std::string& name) {
        const std::string f = lower(name);
        if (f == "exp") return taco_spectral::F_EXP;
        if (f == "log") return taco_spectral::F_LOG;
        if (f == "sqrt") return taco_spectral::F_SQRT;
        if (f == "abs") return taco_spectral::F_ABS;
        if (f == "sin") return taco_spectral::F_SIN;
        if (f == "cos") return taco_spectral::F_COS;
        if (f == "tan") return taco_spectral::F_TAN;
        if (f == "step" || f == "heaviside") return taco_spectral::F_STEP;
        if (f == "pow") return taco_spectral::F_POW;
        if (f == "max") return taco_spectral::F_MAX;
        if (f == "min") return taco_spectral::F_MIN;
        throw Error(TACO_ERR_INVALID, "Unknown function: " + name);
    }

    void run() {
        enum Prev { P_OP, P_VAL, P_LPAREN, P_COMMA, P_FUNC } prev = P_OP;
        std::size_t i = 0;
        while (i < src.size()) {
            const char c = src[i];
            if (std::isspace(static_cast<unsigned char>(c))) {
                ++i;
                continue;
            }
            if (std::isdigit(static_cast<unsigned char>(c)) || c == '.') {
                const char* b = src.c_str() + i;
                char* e = nullptr;
                const double v = std::strtod(b, &e);
                if (e == b) throw Error(TACO_ERR_INVALID, "Invalid number in expression");
                emit(taco_spectral::PUSH, v);
                i += static_cast<std::size_t>(e - b);
                prev = P_VAL;
                continue;
            }
            if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
                std::size_t j = i + 1;
                while (j < src.size() && (std::isalnum(static_cast<unsigned char>(src[j])) || src[j] == '_')) ++j;
                const std::string name = src.substr(i, j - i);
                std::size_t k = j;
                while (k < src.size() && std::isspace(static_cast<unsigned char>(src[k]))) ++k;
                i = j;
                if (k < src.size() && src[k] == '(') {  // a function call
                    st.push_back({Pending::FUNC, func_code(name), 1});
                    prev = P_FUNC;
                    continue;
                }
                const std::string l = lower(name);
                if (l == "w") {
                    emit(taco_spectral::VAR_W);
                } else if (l == "pi") {
                    emit(taco_spectral::PUSH, 3.141592653589793238462643383279502884);
                } else if (l == "e") {
                    emit(taco_spectral::PUSH, 2.718281828459045235360287471352662498);
                } else {
                    auto it = std::find_if(params.begin(), params.end(), [&](const auto& kv) { return kv.first == name; });
                    if (it == params.end())
                        throw Error(TACO_ERR_INVALID, "Unknown identifier (not w/pi/e and not in params): " + name);
                    emit(taco_spectral::PUSH, it->second);
                }
                prev = P_VAL;
                continue;
            }
            ++i;
            if (c == '(') {
                st.push_back({Pending::PAREN, taco_spectral::PUSH, 0});
                prev = P_LPAREN;
            } else if (c == ',') {
                while (!st.empty() && st.back().kind != Pending::PAREN) pop_to_output();
                if (st.size() < 2 || st[st.size() - 2].kind != Pending::FUNC)
                    throw Error(TACO_ERR_INVALID, "Comma outside function call");
                st[st.size() - 2].args += 1;
                prev = P_COMMA;
            } else if (c == ')') {
                while (!st.empty() && st.back().kind != Pending::PAREN) pop_to_output();
                if (st.empty()) throw Error(TACO_ERR_INVALID, "Mismatched ')'");
                st.pop_back();
                if (!st.empty() && st.back().kind == Pending::FUNC) pop_to_output();
                prev = P_VAL;
            } else if (c == '+' || c == '-' || c == '*' || c == '/' || c == '^') {
                const bool unary = c == '-' && (prev == P_OP || prev == P_LPAREN || prev == P_COMMA);
                const Code op = unary ? taco_spectral::NEG
                                      : c == '+'   ? taco_spectral::ADD
                                      : c == '-'   ? taco_spectral::SUB
                                      : c == '*'   ? taco_spectral::MUL
                                      : c == '/'   ? taco_spectral::DIV
                                                   : taco_spectral::POW;
                while (!st.empty() && st.back().kind == Pending::OP) {
                    const int pt = prec(st.back().code), p = prec(op);
                    if (pt > p || (!right(op) && pt == p)) pop_to_output();
                    else break;
                }
                st.push_back({Pending::OP, op, 0});
                prev = P_OP;
            } else {
                throw Error(TACO_ERR_INVALID, std::string("Unexpected character in expression: '") + c + "'");
            }
        }
        while (!st.empty()) {
            if (st.back().kind == Pending::PAREN) throw Error(TACO_ERR_INVALID, "Mismatched '('");
            pop_to_output();
        }
        // the program must leave exactly one value, never reading below the stack
        if (out.empty()) throw Error(TACO_ERR_INVALID, "Expression is empty");
        long depth = 0;
        for (const auto& in : out) {
            const int a = arity_of(in.code);
            if (depth < a) throw Error(TACO_ERR_INVALID, "Expression stack underflow");
            depth += 1 - a;
        }
        if (depth != 1) throw Error(TACO_ERR_INVALID, "Expression did not reduce to a scalar");
    }
};

}  // namespace

namespace taco_engine {

// Rectangle-rule prefix integrals (gamma.hpp:127-160):
// G[k] = sum_{m <= k} (dt phi_m) C[m], phi_m = step^m, step = exp(i w dt).
void gamma_rect(const double* Cin, std::size_t N, double dt, const double* omegas, int M, double* Gout) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "gamma_rect: dt must be > 0");
    if (N == 0 || M <= 0) return;
    using cd = std::complex<double>;
    const cd* C = reinterpret_cast<const cd*>(Cin);
    cd* G = reinterpret_cast<cd*>(Gout);
#pragma omp parallel for schedule(static) if (N * static_cast<std::size_t>(M) >= 200000)
    for (int j = 0; j < M; ++j) {
        cd* col = G + static_cast<std::size_t>(j) * N;
        const cd step = expi_in_mode(omegas[j] * dt);
        cd phi(1.0, 0.0);
        cd acc = (dt * phi) * C[0];
        col[0] = acc;
        for (std::size_t k = 1; k < N; ++k) {
            phi *= step;
            acc += (dt * phi) * C[k];
            col[k] = acc;
        }
    }
}

}  // namespace taco_engine

extern "C" {

int taco_spectral_create(const char* source, const char* const* names, const double* values, int nparams,
                         taco_spectral** out) {
    return taco_engine::guarded([&] {
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (!source) throw Error(TACO_ERR_INVALID, "null expression");
        if (nparams < 0 || (nparams > 0 && (!names || !values))) throw Error(TACO_ERR_INVALID, "bad parameter list");
        std::vector<std::pair<std::string, double>> params;
        for (int q = 0; q < nparams; ++q) {
            if (!names[q]) throw Error(TACO_ERR_INVALID, "null parameter name");
            params.emplace_back(names[q], values[q]);
        }
        auto J = std::make_unique<taco_spectral>();
        J->source = source;
        Compiler comp{J->source, params, {}, {}};
        comp.run();
        J->prog = std::move(comp.out);
        *out = J.release();
    });
}

int taco_spectral_eval(const taco_spectral* J, const double* w, size_t n, double* out) {
    return taco_engine::guarded([&] {
        if (!J) throw Error(TACO_ERR_INVALID, "null taco_spectral");
        if (n && (!w || !out)) throw Error(TACO_ERR_INVALID, "null array");
        for (size_t k = 0; k < n; ++k) out[k] = J->eval(w[k]);
    });
}

int taco_spectral_destroy(taco_spectral* J) {
    delete J;
    return TACO_OK;
}

int taco_bcf_fft_spectral(size_t N, double dt, const taco_spectral* J, double beta, double* C_out) {
    return taco_engine::guarded([&] {
        if (!J) throw Error(TACO_ERR_INVALID, "null taco_spectral");
        taco_engine::bcf_fft_fn(N, dt, [J](double w) { return J->eval(w); }, beta, C_out);
    });
}

int taco_gamma_rect(const double* C, size_t Nc, double dt, const double* omegas, int nf, double* gamma_out) {
    return taco_engine::guarded([&] { taco_engine::gamma_rect(C, Nc, dt, omegas, nf, gamma_out); });
}

}  // extern "C"
