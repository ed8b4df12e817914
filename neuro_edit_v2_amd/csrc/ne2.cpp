// ne2.cpp -- .ne2 network files into the simulator layout.
//
// File: magic "NE2\0", u32 little-endian format version 1..7, then the
// bincode (v1 defaults: little-endian, fixed-width integers, u64 lengths,
// u32 enum variant indices, u8 Option tags) encoding of `Network` in that
// version's field layout (net-format/src/io.rs:11-12,437-468; older layouts
// and their lifts: io.rs:14-330).  The decoded network is validated
// (network.rs:1007-1150, the checks that bear on simulation) and turned
// into the flat arrays the step kernels take:
//   NeuronParams::from_network  (sim/src/state.rs:98-153)
//   Network::synapse_csr(dt)    (net-format/src/network.rs:654-692)
//   StimMap::from_network       (sim/src/stim.rs:34-57)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/neuro_hip.h"
#include "host_util.h"
#include "netgen.h"

using nh::set_error;

namespace {

// ---------------------------------------------------------------- decoding
struct Cur {
    const uint8_t *p, *e;
    std::string err;
    bool need(uint64_t k) {
        if (!err.empty()) return false;
        if ((uint64_t)(e - p) < k) {
            err = "truncated file";
            return false;
        }
        return true;
    }
    template <class T> T pod() {
        T v{};
        if (need(sizeof(T))) {
            std::memcpy(&v, p, sizeof(T));
            p += sizeof(T);
        }
        return v;
    }
    uint8_t u8() { return pod<uint8_t>(); }
    uint16_t u16() { return pod<uint16_t>(); }
    uint32_t u32() { return pod<uint32_t>(); }
    uint64_t u64() { return pod<uint64_t>(); }
    float f32() { return pod<float>(); }
    // sequence length; bounded by the bytes left (every element takes >= min_elem)
    uint64_t len(uint64_t min_elem) {
        const uint64_t n = u64();
        if (err.empty() && min_elem && n > (uint64_t)(e - p) / min_elem) err = "length exceeds the file";
        return err.empty() ? n : 0;
    }
    std::string str() {
        const uint64_t n = len(1);
        std::string s;
        if (need(n)) {
            s.assign((const char *)p, n);
            p += n;
        }
        return s;
    }
    std::vector<float> vf32() {
        const uint64_t n = len(4);
        std::vector<float> v(n);
        for (auto &x : v) x = f32();
        return v;
    }
    std::vector<uint32_t> vu32() {
        const uint64_t n = len(4);
        std::vector<uint32_t> v(n);
        for (auto &x : v) x = u32();
        return v;
    }
    bool opt() {
        const uint8_t t = u8();
        if (t > 1 && err.empty()) err = "bad Option tag";
        return t == 1;
    }
    void f32s(float *out, int k) {
        for (int q = 0; q < k; ++q) out[q] = f32();
    }
};

enum Kind : uint32_t { kExc = 0, kInh = 1, kRegular = 2 };

struct Type {
    std::string name, model;
    std::vector<float> params;
    uint32_t kind = kExc;
};
struct Neuron {
    std::string name;
    float pos[3] = {0, 0, 0};
    uint32_t ty = 0;
    bool has_params = false;
    std::vector<float> params;
};
struct Syn {
    uint32_t src, tgt;
    float w, delay_ms, tau_ms;
};
struct Grid {
    uint32_t dims[3];
    float origin[3], spacing;
    uint8_t axes[3];
};
struct Group {
    std::string name;
    std::vector<uint32_t> members;
    bool has_grid = false;
    Grid grid{};
};
struct Input {
    std::string name;
    std::vector<std::pair<uint32_t, float>> targets;
    float pos[3];
};
struct Readout {
    std::string name;
    std::vector<uint32_t> units;
    std::vector<float> weights;
    float tau_ms = 1.0f;
};
struct Meta {
    float dt_ms = 0.1f, conduction_speed = 1.0f, synapse_tau_ms = 2.0f, readout_tau_ms = 2.0f;
    float noise_rate_hz = 0.0f, noise_kick_mv = 20.0f;
    uint64_t seed = 1;
};
struct Net {
    std::vector<Type> types;
    std::vector<Neuron> neurons;
    std::vector<Syn> syn;
    std::vector<Group> groups;
    std::vector<Input> inputs;
    uint64_t n_samples = 0, n_scripts = 0;
    std::vector<Readout> readouts;
    Meta meta;
};

uint32_t kind_of(Cur &c) {
    const uint32_t k = c.u32();
    if (k > kRegular && c.err.empty()) c.err = "bad NeuronKind";
    return k;
}

// NeuronType: name, model, params, kind [, color: Option<[f32;3]> from v5]
void de_types(Cur &c, Net &n, bool with_color) {
    const uint64_t k = c.len(8 + 8 + 8 + 4);
    n.types.resize(k);
    for (auto &t : n.types) {
        t.name = c.str();
        t.model = c.str();
        t.params = c.vf32();
        t.kind = kind_of(c);
        if (with_color && c.opt()) {
            float col[3];
            c.f32s(col, 3);
        }
    }
}

// Neuron (v4+): name, pos, ty u16, params Option<Box<[f32]>>
void de_neurons(Cur &c, Net &n) {
    const uint64_t k = c.len(8 + 12 + 2 + 1);
    n.neurons.resize(k);
    for (auto &x : n.neurons) {
        x.name = c.str();
        c.f32s(x.pos, 3);
        x.ty = c.u16();
        x.has_params = c.opt();
        if (x.has_params) x.params = c.vf32();
    }
}

// NeuronV3 (v1-v3): name, pos, model (enum: 0 = Izhikevich{a,b,c,d}), kind;
// lifted into a type table (io.rs:98-170): one type per distinct (params
// bits, kind) in first-use order, named exc / inh / regular for the stock
// rows or izh#xxxx (low 16 bits of an FNV-1a over the param bytes and the
// kind), a clashing name taking the first free suffix 2, 3, ...
void de_lift_neurons(Cur &c, Net &n) {
    static const float RS[4] = {0.02f, 0.2f, -65.0f, 8.0f}, FS[4] = {0.1f, 0.2f, -65.0f, 2.0f};
    const uint64_t k = c.len(8 + 12 + 4 + 16 + 4);
    n.neurons.resize(k);
    std::map<std::pair<std::vector<uint32_t>, uint32_t>, uint32_t> lut;
    for (auto &x : n.neurons) {
        x.name = c.str();
        c.f32s(x.pos, 3);
        const uint32_t variant = c.u32();
        if (variant != 0 && c.err.empty()) c.err = "bad NeuronModel";
        float p[4];
        c.f32s(p, 4);
        const uint32_t kind = kind_of(c);
        if (!c.err.empty()) return;
        std::vector<uint32_t> bits(4);
        for (int q = 0; q < 4; ++q) std::memcpy(&bits[q], &p[q], 4);
        auto key = std::make_pair(bits, kind);
        auto it = lut.find(key);
        if (it == lut.end()) {
            auto same = [&](const float *row) { return std::memcmp(row, p, 16) == 0; };
            std::string name;
            if (same(RS) && kind == kExc) name = "exc";
            else if (same(FS) && kind == kInh) name = "inh";
            else if (same(RS) && kind == kRegular) name = "regular";
            else {
                uint64_t h = 0xcbf29ce484222325ull;
                for (int q = 0; q < 4; ++q)
                    for (int b = 0; b < 4; ++b) h = (h ^ ((bits[q] >> (8 * b)) & 0xFF)) * 0x100000001b3ull;
                h = (h ^ kind) * 0x100000001b3ull;
                char buf[16];
                std::snprintf(buf, sizeof buf, "izh#%04x", (unsigned)(h & 0xffff));
                name = buf;
            }
            const std::string base = name;
            for (uint32_t s = 2;; ++s) {
                bool used = false;
                for (auto &t : n.types) used = used || t.name == name;
                if (!used) break;
                name = base + std::to_string(s);
            }
            Type t;
            t.name = name;
            t.model = "izh";
            t.params.assign(p, p + 4);
            t.kind = kind;
            n.types.push_back(t);
            it = lut.emplace(key, (uint32_t)n.types.size() - 1).first;
        }
        x.ty = it->second;
    }
}

// Synapse: src, tgt, w, delay_ms [, tau_ms from v3; older lift to 0]
void de_syn(Cur &c, Net &n, bool with_tau) {
    const uint64_t k = c.len(with_tau ? 20 : 16);
    n.syn.resize(k);
    for (auto &s : n.syn) {
        s.src = c.u32();
        s.tgt = c.u32();
        s.w = c.f32();
        s.delay_ms = c.f32();
        s.tau_ms = with_tau ? c.f32() : 0.0f;
    }
}

// Group: name, members [, grid: Option<GridShape> from v2]
void de_groups(Cur &c, Net &n, bool with_grid) {
    const uint64_t k = c.len(16);
    n.groups.resize(k);
    for (auto &g : n.groups) {
        g.name = c.str();
        g.members = c.vu32();
        if (with_grid && c.opt()) {
            g.has_grid = true;
            for (auto &d : g.grid.dims) d = c.u32();
            c.f32s(g.grid.origin, 3);
            g.grid.spacing = c.f32();
            for (auto &a : g.grid.axes) a = c.u8();
        }
    }
}

// Input: name, targets Vec<(u32, f32)>, pos
void de_inputs(Cur &c, Net &n) {
    const uint64_t k = c.len(8 + 8 + 12);
    n.inputs.resize(k);
    for (auto &in : n.inputs) {
        in.name = c.str();
        const uint64_t m = c.len(8);
        in.targets.resize(m);
        for (auto &t : in.targets) {
            t.first = c.u32();
            t.second = c.f32();
        }
        c.f32s(in.pos, 3);
    }
}

// Samples: name, times, values (not used by the simulator)
void de_samples(Cur &c, Net &n) {
    n.n_samples = c.len(24);
    for (uint64_t q = 0; q < n.n_samples && c.err.empty(); ++q) {
        (void)c.str();
        (void)c.vf32();
        (void)c.vf32();
    }
}

// MlpDef: sizes, w Vec<Vec<f32>>, mu, sd [, act, head: u32 enums from v7]
void de_mlp(Cur &c, bool v7) {
    (void)c.vu32();
    const uint64_t k = c.len(8);
    for (uint64_t q = 0; q < k && c.err.empty(); ++q) (void)c.vf32();
    (void)c.vf32();
    (void)c.vf32();
    if (v7) {
        if (c.u32() > 1 && c.err.empty()) c.err = "bad MlpAct";
        if (c.u32() > 1 && c.err.empty()) c.err = "bad MlpHead";
    }
}

// Readout: name, units, weights, bias, bias_weight, tau_ms, pos
//          [, mlp: Option<MlpDef> from v6] [, mlp_layers from v7]
void de_readouts(Cur &c, Net &n, int version) {
    const uint64_t k = c.len(8 + 8 + 8 + 24);
    n.readouts.resize(k);
    for (auto &r : n.readouts) {
        r.name = c.str();
        r.units = c.vu32();
        r.weights = c.vf32();
        (void)c.f32();  // bias
        (void)c.f32();  // bias_weight
        r.tau_ms = c.f32();
        float pos[3];
        c.f32s(pos, 3);
        if (version >= 6 && c.opt()) de_mlp(c, version >= 7);
        if (version >= 7) (void)c.vu32();
    }
}

void de_scripts(Cur &c, Net &n) {
    n.n_scripts = c.len(16);
    for (uint64_t q = 0; q < n.n_scripts && c.err.empty(); ++q) {
        (void)c.str();
        (void)c.str();
    }
}

// Meta: dt_ms, seed, conduction_speed, synapse_tau_ms, readout_tau_ms,
// noise_rate_hz, noise_kick_mv
void de_meta(Cur &c, Net &n) {
    Meta &m = n.meta;
    m.dt_ms = c.f32();
    m.seed = c.u64();
    m.conduction_speed = c.f32();
    m.synapse_tau_ms = c.f32();
    m.readout_tau_ms = c.f32();
    m.noise_rate_hz = c.f32();
    m.noise_kick_mv = c.f32();
}

int decode(const uint8_t *b, uint64_t len, Net &n, uint32_t &version) {
    if (len < 8 || std::memcmp(b, "NE2\0", 4) != 0) return set_error(NH_ERR_NET, "not a .ne2 file (bad magic)");
    std::memcpy(&version, b + 4, 4);
    if (version < 1 || version > 7)
        return set_error(NH_ERR_NET, "unsupported .ne2 version " + std::to_string(version));
    Cur c{b + 8, b + len, {}};
    if (version <= 3) {
        de_lift_neurons(c, n);
        de_syn(c, n, version >= 3);
        de_groups(c, n, version >= 2);
    } else {
        de_types(c, n, version >= 5);
        de_neurons(c, n);
        de_syn(c, n, true);
        de_groups(c, n, true);
    }
    de_inputs(c, n);
    de_samples(c, n);
    de_readouts(c, n, version);
    de_scripts(c, n);
    de_meta(c, n);
    if (!c.err.empty()) return set_error(NH_ERR_NET, "decode error: " + c.err);
    return NH_OK;
}

// ---------------------------------------------------------------- validate
// The checks of Network::validate (network.rs:1007-1150) that bear on the
// simulator; readout MLP shapes are not checked (readouts are not run here).
int validate(const Net &n) {
    const uint32_t N = (uint32_t)n.neurons.size();
    auto fin = [](float x) { return std::isfinite(x); };
    const Meta &m = n.meta;
    if (!(fin(m.dt_ms) && m.dt_ms > 0.0f)) return set_error(NH_ERR_NET, "meta.dt_ms must be positive and finite");
    if (!(fin(m.conduction_speed) && m.conduction_speed > 0.0f))
        return set_error(NH_ERR_NET, "meta.conduction_speed must be positive and finite");
    for (float v : {m.synapse_tau_ms, m.readout_tau_ms, m.noise_rate_hz, m.noise_kick_mv})
        if (!(fin(v) && v >= 0.0f)) return set_error(NH_ERR_NET, "meta rates and taus must be finite and >= 0");
    for (size_t i = 0; i < n.types.size(); ++i) {
        const Type &t = n.types[i];
        if (t.name.empty()) return set_error(NH_ERR_NET, "type with an empty name");
        for (size_t j = 0; j < i; ++j)
            if (n.types[j].name == t.name) return set_error(NH_ERR_NET, "duplicate type name " + t.name);
        if (t.model.empty()) return set_error(NH_ERR_NET, "type " + t.name + " has an empty model id");
        if (t.params.empty()) return set_error(NH_ERR_NET, "type " + t.name + " has bad params");
        for (float p : t.params)
            if (!fin(p)) return set_error(NH_ERR_NET, "type " + t.name + " has bad params");
    }
    if (n.types.size() > 65536) return set_error(NH_ERR_NET, "types exceed the u16 ty index");
    for (uint32_t i = 0; i < N; ++i) {
        const Neuron &x = n.neurons[i];
        for (float p : x.pos)
            if (!fin(p)) return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + " has non-finite pos");
        if (x.ty >= n.types.size()) return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + " ty out of range");
        if (x.has_params) {
            if (x.params.size() != n.types[x.ty].params.size())
                return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + " has a bad param override");
            for (float p : x.params)
                if (!fin(p)) return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + " has a bad param override");
        }
    }
    for (size_t i = 0; i < n.syn.size(); ++i) {
        const Syn &s = n.syn[i];
        const std::string id = "synapse " + std::to_string(i);
        if (s.src >= N || s.tgt >= N) return set_error(NH_ERR_NET, id + " references a missing neuron");
        if (!(s.tau_ms >= 0.0f && fin(s.tau_ms))) return set_error(NH_ERR_NET, id + " has bad tau_ms");
        if (!fin(s.w)) return set_error(NH_ERR_NET, id + " has a non-finite weight");
        if (!(s.delay_ms >= 0.0f && fin(s.delay_ms))) return set_error(NH_ERR_NET, id + " has bad delay_ms");
    }
    for (const Group &g : n.groups) {
        for (uint32_t mb : g.members)
            if (mb >= N) return set_error(NH_ERR_NET, "group '" + g.name + "' references a missing neuron");
        if (g.has_grid) {
            bool seen[3] = {false, false, false};
            for (uint8_t a : g.grid.axes) {
                if (a >= 3 || seen[a]) return set_error(NH_ERR_NET, "group '" + g.name + "' has invalid grid axes");
                seen[a] = true;
            }
            if (!(fin(g.grid.spacing) && g.grid.spacing != 0.0f))
                return set_error(NH_ERR_NET, "group '" + g.name + "' has a bad grid spacing");
            for (uint32_t d : g.grid.dims)
                if (d == 0) return set_error(NH_ERR_NET, "group '" + g.name + "' grid dims contain a zero");
        }
    }
    for (const Input &in : n.inputs)
        for (auto &t : in.targets) {
            if (t.first >= N) return set_error(NH_ERR_NET, "input '" + in.name + "' targets a missing neuron");
            if (!fin(t.second)) return set_error(NH_ERR_NET, "input '" + in.name + "' has a non-finite weight");
        }
    for (const Readout &r : n.readouts) {
        for (uint32_t u : r.units)
            if (u >= N) return set_error(NH_ERR_NET, "readout '" + r.name + "' references a missing unit");
        if (!r.weights.empty() && r.weights.size() != r.units.size())
            return set_error(NH_ERR_NET, "readout '" + r.name + "' weights do not match its units");
        if (!(fin(r.tau_ms) && r.tau_ms > 0.0f))
            return set_error(NH_ERR_NET, "readout '" + r.name + "' tau_ms must be positive and finite");
    }
    return NH_OK;
}

// ---------------------------------------------------------------- to the simulator layout
int to_sim(const Net &n, float dt, nh_net &out) {
    const uint32_t N = (uint32_t)n.neurons.size();
    out.n = N;
    // NeuronParams::from_network: model table in first-use order (registry
    // ids: 0 izh with 4 params, 1 adex with 10), rows padded to the widest
    out.model_ids.clear();
    out.model.assign(N, 0);
    uint32_t width = 0;
    std::vector<const std::vector<float> *> rows(N);
    std::vector<uint8_t> reg(N);
    for (uint32_t i = 0; i < N; ++i) {
        const Neuron &x = n.neurons[i];
        const Type &t = n.types[x.ty];
        uint8_t id;
        uint32_t np;
        if (t.model == "izh") id = 0, np = 4;
        else if (t.model == "adex") id = 1, np = 10;
        else return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + ": unknown model " + t.model);
        const std::vector<float> &p = x.has_params ? x.params : t.params;
        if (p.size() != np)
            return set_error(NH_ERR_NET, "neuron " + std::to_string(i) + ": " + std::to_string(p.size()) +
                                             " params, model " + t.model + " wants " + std::to_string(np));
        uint32_t mi = 0;
        while (mi < out.model_ids.size() && out.model_ids[mi] != id) ++mi;
        if (mi == out.model_ids.size()) {
            if (mi >= 8) return set_error(NH_ERR_NET, "more than 8 models");
            out.model_ids.push_back(id);
        }
        out.model[i] = (uint8_t)mi;
        width = std::max(width, np);
        rows[i] = &p;
        reg[i] = id;
    }
    out.n_params = std::max<uint32_t>(width, 4);
    out.n_state = 2;
    out.p.assign((size_t)N * out.n_params, 0.0f);
    out.x0.assign((size_t)N * 2, 0.0f);
    for (uint32_t i = 0; i < N; ++i) {
        const std::vector<float> &p = *rows[i];
        std::copy(p.begin(), p.end(), out.p.begin() + (size_t)i * out.n_params);
        if (reg[i] == 0) {  // izh.rs:175-178
            out.x0[2 * (size_t)i] = -65.0f;
            out.x0[2 * (size_t)i + 1] = p[1] * -65.0f;
        } else {  // adex.rs:149-152
            out.x0[2 * (size_t)i] = p[2];
            out.x0[2 * (size_t)i + 1] = 0.0f;
        }
    }
    // synapse_csr: stable sort by source, Dale sign from the source's kind,
    // delay steps round(delay/dt) (half away from zero) at least 1
    out.src_ptr.assign(N + 1, 0);
    for (const Syn &s : n.syn) out.src_ptr[s.src + 1]++;
    for (uint32_t i = 0; i < N; ++i) out.src_ptr[i + 1] += out.src_ptr[i];
    const size_t E = n.syn.size();
    out.tgt.assign(E, 0);
    out.w.assign(E, 0.0f);
    out.delay.assign(E, 0);
    out.tau.assign(E, 0.0f);
    std::vector<uint32_t> fill(out.src_ptr.begin(), out.src_ptr.end() - 1);
    bool any_tau = false;
    for (const Syn &s : n.syn) {
        const uint32_t j = fill[s.src]++;
        const bool inh = n.types[n.neurons[s.src].ty].kind == kInh;
        out.tgt[j] = s.tgt;
        out.w[j] = inh ? -std::fabs(s.w) : std::fabs(s.w);
        out.tau[j] = s.tau_ms;
        any_tau = any_tau || s.tau_ms != 0.0f;
        float steps = std::round(s.delay_ms / dt);
        if (steps < 1.0f) steps = 1.0f;
        if (!std::isfinite(steps) || steps > 65535.0f)
            return set_error(NH_ERR_NET, "synapse " + std::to_string(s.src) + " -> " + std::to_string(s.tgt) +
                                             ": delay exceeds the 65535-step ring");
        out.delay[j] = (uint16_t)steps;
    }
    if (!any_tau) out.tau.clear();
    // StimMap::from_network: per neuron, (input, weight) in input order
    out.n_inputs = (uint32_t)n.inputs.size();
    std::vector<std::vector<std::pair<uint32_t, float>>> per(N);
    for (uint32_t k = 0; k < n.inputs.size(); ++k)
        for (auto &t : n.inputs[k].targets) per[t.first].push_back({k, t.second});
    out.stim_ptr.assign(N + 1, 0);
    out.stim_input.clear();
    out.stim_w.clear();
    for (uint32_t i = 0; i < N; ++i) {
        out.stim_ptr[i + 1] = out.stim_ptr[i] + (uint32_t)per[i].size();
        for (auto &e : per[i]) {
            out.stim_input.push_back(e.first);
            out.stim_w.push_back(e.second);
        }
    }
    // editor-level tables the simulator's callers address by index
    out.type_names.clear();
    for (const Type &t : n.types) out.type_names.push_back(t.name);
    out.neuron_type.resize(N);
    for (uint32_t i = 0; i < N; ++i) out.neuron_type[i] = (uint16_t)n.neurons[i].ty;
    out.input_names.clear();
    for (const Input &in : n.inputs) out.input_names.push_back(in.name);
    out.groups.clear();
    for (const Group &g : n.groups) out.groups.push_back({g.name, g.members});
    return NH_OK;
}

}  // namespace

extern "C" {

int nh_net_load_ne2(const uint8_t *bytes, uint64_t len, float dt_ms, nh_net **out, nh_ne2_info *info) {
    return nh::guarded([&]() -> int {
        if (!bytes || !out) return set_error(NH_ERR_ARG, "null argument");
        *out = nullptr;
        Net n;
        uint32_t version = 0;
        int rc = decode(bytes, len, n, version);
        if (rc) return rc;
        if ((rc = validate(n))) return rc;
        const float dt = dt_ms > 0.0f ? dt_ms : n.meta.dt_ms;
        std::unique_ptr<nh_net> net(new (std::nothrow) nh_net());
        if (!net) return set_error(NH_ERR_CAPACITY, "out of host memory");
        if ((rc = to_sim(n, dt, *net))) return rc;
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->version = version;
            info->dt_ms = n.meta.dt_ms;
            info->seed = n.meta.seed;
            info->conduction_speed = n.meta.conduction_speed;
            info->synapse_tau_ms = n.meta.synapse_tau_ms;
            info->readout_tau_ms = n.meta.readout_tau_ms;
            info->noise_rate_hz = n.meta.noise_rate_hz;
            info->noise_kick_mv = n.meta.noise_kick_mv;
            info->n_types = (uint32_t)n.types.size();
            info->n_groups = (uint32_t)n.groups.size();
            info->n_inputs = (uint32_t)n.inputs.size();
            info->n_samples = (uint32_t)n.n_samples;
            info->n_readouts = (uint32_t)n.readouts.size();
            info->n_scripts = (uint32_t)n.n_scripts;
        }
        *out = net.release();
        return NH_OK;
    });
}

int nh_net_type_name(const nh_net *net, uint32_t k, char *buf, uint32_t cap) {
    if (!net || !buf || !cap) return set_error(NH_ERR_ARG, "null argument");
    if (k >= net->type_names.size()) return set_error(NH_ERR_ARG, "type index out of range");
    std::strncpy(buf, net->type_names[k].c_str(), cap - 1);
    buf[cap - 1] = 0;
    return NH_OK;
}

int nh_net_neuron_types(const nh_net *net, uint16_t *ty, uint32_t n) {
    if (!net || (!ty && n)) return set_error(NH_ERR_ARG, "null argument");
    if (n < net->neuron_type.size()) return set_error(NH_ERR_ARG, "output too small");
    std::copy(net->neuron_type.begin(), net->neuron_type.end(), ty);
    return NH_OK;
}

int nh_net_input_name(const nh_net *net, uint32_t k, char *buf, uint32_t cap) {
    if (!net || !buf || !cap) return set_error(NH_ERR_ARG, "null argument");
    if (k >= net->input_names.size()) return set_error(NH_ERR_ARG, "input index out of range");
    std::strncpy(buf, net->input_names[k].c_str(), cap - 1);
    buf[cap - 1] = 0;
    return NH_OK;
}

int nh_net_group(const nh_net *net, uint32_t k, char *name, uint32_t name_cap, uint32_t *members,
                 uint32_t cap, uint32_t *n_members) {
    if (!net || !n_members) return set_error(NH_ERR_ARG, "null argument");
    if (k >= net->groups.size()) return set_error(NH_ERR_ARG, "group index out of range");
    const auto &g = net->groups[k];
    *n_members = (uint32_t)g.second.size();
    if (name && name_cap) {
        std::strncpy(name, g.first.c_str(), name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (members) {
        if (cap < g.second.size()) return set_error(NH_ERR_CAPACITY, "members array too small");
        std::copy(g.second.begin(), g.second.end(), members);
    }
    return NH_OK;
}

}  // extern "C"
