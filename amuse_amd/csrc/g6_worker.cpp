// Community-code worker for the device-resident block-step integrator
// (g6d_* in include/g6x.h), speaking AMUSE's socket message protocol so that
// an AMUSE-style Python proxy can drive it out of process.
//
// Protocol (restated from the Python side, src/amuse/rfi/channel.py:1720-1954
// SocketMessage, and the generated C worker, src/amuse/rfi/tools/create_c.py:
// constants :52-66, socket loop :812-960, handle_call :1334-1360, argument
// packing :1081-1210):
//   * the worker is started as `g6_worker PORT HOST false` and connects to
//     HOST:PORT (TCP, no delay);
//   * a message is a header of 11 int32 — flags (4 bytes: big-endian,
//     error, has-units, unused), call id, function id, call count, and the
//     number of ints, longs, floats, doubles, booleans, strings and units —
//     followed by the int32, int64, float32, float64, bool (1 byte) arrays,
//     the strings (int32 lengths, then the characters, each string followed
//     by a NUL) and the units (float64);
//   * function id = |crc32(name)| with crc32 read as a signed int32
//     (src/amuse/rfi/core.py:368-417); id 0 stops the worker after its reply;
//   * the parameters of one dtype are numbered in order among the inputs and
//     among the outputs; input k of call i is in[k * call_count + i]; the
//     int32 result is output 0 of the ints, so int outputs start at 1;
//   * an unknown id answers with the error flag (flags word 256) and one
//     string;
//   * a header is checked before anything is allocated: a call count above
//     kMaxCallCount, a negative length, or arrays that would hold more than
//     kMaxMessageBytes in all, end the session (the stream cannot be
//     re-synchronised), as a closed channel does.
//
// Functions: ph4's interface (src/amuse_ph4/interface.cc, interface.py,
// src/amuse/community/interface/gd.py) with the same names, parameters and
// return codes, backed by the device-resident driver.  After
// commit_particles, synchronize_model steps every particle to the system
// time (ph4's synchronize_all); particles added, deleted or changed
// (new_particle, delete_particle, set_state/mass/position/velocity) edit the
// host copy of the state, and recommit_particles (or the next evolve_model)
// rebuilds the driver from it at the system time: forces and initial steps
// recomputed, as ph4's recommit_particles does (interface.cc:349-379), with
// the model time and the step counters carried over.
#include "../../include/g6x.h"

#include <arpa/inet.h>
#include <netdb.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// function ids: CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320, initial
// and final value 0xFFFFFFFF) of the name, read as signed, absolute value
// ---------------------------------------------------------------------------
uint32_t crc32_ieee(const char *s)
{
    uint32_t c = 0xFFFFFFFFu;
    for (; *s; s++) {
        c ^= static_cast<uint8_t>(*s);
        for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c ^ 0xFFFFFFFFu;
}

int function_id(const char *name)
{
    const int64_t x = static_cast<int64_t>(crc32_ieee(name));
    const int64_t signed_x = x - ((x & 0x80000000LL) << 1);
    return static_cast<int>(signed_x < 0 ? -signed_x : signed_x);
}

// ---------------------------------------------------------------------------
// messages
// ---------------------------------------------------------------------------
enum { H_FLAGS, H_CALL_ID, H_FUNCTION_ID, H_CALL_COUNT, H_INTS, H_LONGS, H_FLOATS, H_DOUBLES,
       H_BOOLS, H_STRINGS, H_UNITS, H_SIZE };
constexpr int kErrorFlag = 256;
// Limits on what one message may declare (AMUSE splits long vectorised calls
// into messages of at most 1,000,000 calls, src/amuse/rfi/channel.py
// max_message_length; ph4's new_particle for 2^20 particles is ~80 MB).
constexpr int32_t kMaxCallCount = 1 << 24;
constexpr size_t kMaxMessageBytes = size_t(1) << 30;

struct Message {
    int32_t header[H_SIZE] = {};
    std::vector<int32_t> ints;
    std::vector<int64_t> longs;
    std::vector<float> floats;
    std::vector<double> doubles;
    std::vector<uint8_t> bools;
    std::vector<std::string> strings;
    std::vector<double> units;
};

bool recv_all(int fd, void *p, size_t n)
{
    char *c = static_cast<char *>(p);
    while (n > 0) {
        const ssize_t r = recv(fd, c, n, 0);
        if (r <= 0) return false;
        c += r;
        n -= size_t(r);
    }
    return true;
}

bool send_all(int fd, const void *p, size_t n)
{
    const char *c = static_cast<const char *>(p);
    while (n > 0) {
        const ssize_t r = send(fd, c, n, 0);
        if (r <= 0) return false;
        c += r;
        n -= size_t(r);
    }
    return true;
}

template <class T>
bool recv_vec(int fd, std::vector<T> &v, int n)
{
    if (n < 0) return false;
    v.resize(size_t(n));
    return n == 0 || recv_all(fd, v.data(), sizeof(T) * size_t(n));
}

// Bytes the header's array lengths declare (strings without their
// characters, which are bounded once their lengths are read), or SIZE_MAX
// when a length is negative or the call count is out of range.
size_t declared_bytes(const int32_t *h)
{
    if (h[H_CALL_COUNT] < 0 || h[H_CALL_COUNT] > kMaxCallCount) return SIZE_MAX;
    const int32_t n[] = {h[H_INTS], h[H_LONGS], h[H_FLOATS], h[H_DOUBLES], h[H_BOOLS],
                         h[H_STRINGS], h[H_UNITS]};
    const size_t w[] = {4, 8, 4, 8, 1, 4, 8};
    size_t total = 0;
    for (int k = 0; k < 7; k++) {
        if (n[k] < 0) return SIZE_MAX;
        total += w[k] * size_t(n[k]);
    }
    return total;
}

template <class T>
bool send_vec(int fd, const std::vector<T> &v)
{
    return v.empty() || send_all(fd, v.data(), sizeof(T) * v.size());
}

bool receive(int fd, Message &m)
{
    if (!recv_all(fd, m.header, sizeof(m.header))) return false;
    const uint8_t *flags = reinterpret_cast<const uint8_t *>(&m.header[H_FLAGS]);
    if (flags[0] != 0) {  // big-endian sender: this worker is little-endian
        fprintf(stderr, "g6_worker: message endianness does not match\n");
        return false;
    }
    size_t budget = declared_bytes(m.header);
    if (budget > kMaxMessageBytes) {
        fprintf(stderr, "g6_worker: message header declares invalid or oversized arrays\n");
        return false;
    }
    if (!recv_vec(fd, m.ints, m.header[H_INTS]) || !recv_vec(fd, m.longs, m.header[H_LONGS]) ||
        !recv_vec(fd, m.floats, m.header[H_FLOATS]) ||
        !recv_vec(fd, m.doubles, m.header[H_DOUBLES]) ||
        !recv_vec(fd, m.bools, m.header[H_BOOLS]))
        return false;
    const int ns = m.header[H_STRINGS];
    std::vector<int32_t> len;
    if (!recv_vec(fd, len, ns)) return false;
    size_t total = 0;
    for (int32_t l : len) {
        if (l < 0) return false;
        total += size_t(l) + 1;
    }
    if (total > kMaxMessageBytes - budget) {
        fprintf(stderr, "g6_worker: message strings exceed the size limit\n");
        return false;
    }
    std::vector<char> chars(total);
    if (total && !recv_all(fd, chars.data(), total)) return false;
    m.strings.clear();
    size_t at = 0;
    for (int32_t l : len) {
        m.strings.emplace_back(chars.data() + at, size_t(l));
        at += size_t(l) + 1;
    }
    return recv_vec(fd, m.units, m.header[H_UNITS]);
}

bool reply(int fd, Message &m)
{
    m.header[H_INTS] = int32_t(m.ints.size());
    m.header[H_LONGS] = int32_t(m.longs.size());
    m.header[H_FLOATS] = int32_t(m.floats.size());
    m.header[H_DOUBLES] = int32_t(m.doubles.size());
    m.header[H_BOOLS] = int32_t(m.bools.size());
    m.header[H_STRINGS] = int32_t(m.strings.size());
    m.header[H_UNITS] = int32_t(m.units.size());
    if (!send_all(fd, m.header, sizeof(m.header)) || !send_vec(fd, m.ints) ||
        !send_vec(fd, m.longs) || !send_vec(fd, m.floats) || !send_vec(fd, m.doubles) ||
        !send_vec(fd, m.bools))
        return false;
    if (!m.strings.empty()) {
        std::vector<int32_t> len;
        std::string chars;
        for (const std::string &s : m.strings) {
            len.push_back(int32_t(s.size()));
            chars += s;
            chars.push_back('\0');
        }
        if (!send_vec(fd, len) || !send_all(fd, chars.data(), chars.size())) return false;
    }
    return send_vec(fd, m.units);
}

// ---------------------------------------------------------------------------
// the code: ph4's interface over the device-resident driver
// ---------------------------------------------------------------------------
struct Code {
    // parameters (ph4 defaults: src/amuse_ph4/interface.py:506-628,
    // interface.cc:29-31)
    double eps2 = 0.0, eta = 0.14, begin_time = 0.0;
    double dt_fac = 0.0625, dt_limit = 0.03125, dt_median = 8.0;
    int use_gpu = 1, gpu_id = -1, precision = G6X_FP64;
    // particles in order of addition
    std::vector<int> id;
    std::vector<double> mass, radius, pos, vel;  // pos, vel: [n][3]
    std::map<int, int> slot;                      // id -> index
    int next_id = 1;
    void *drv = nullptr;
    bool state_fresh = false;
    std::vector<double> t, dt;
    // a driver rebuilt by recommit starts at system time 0: the model time
    // is begin_time + time_offset + its system time, the step counters add
    bool dirty = false;  // the host copy was edited after commit
    double time_offset = 0.0;
    long long blocks_offset = 0, steps_offset = 0;

    void reset()
    {
        if (drv) g6d_free(drv);
        drv = nullptr;
        id.clear(); mass.clear(); radius.clear(); pos.clear(); vel.clear(); slot.clear();
        next_id = 1;
        state_fresh = false;
        dirty = false;
        time_offset = 0.0;
        blocks_offset = steps_offset = 0;
    }
    void reindex()
    {
        slot.clear();
        for (size_t q = 0; q < id.size(); q++) slot[id[q]] = int(q);
    }
    int counters(long long &b, long long &p, double &ts)
    {
        b = p = 0;
        ts = 0.0;
        if (drv && g6d_counters(drv, &b, &p, &ts) < 0) return -1;
        b += blocks_offset;
        p += steps_offset;
        ts += time_offset;
        return 0;
    }
    int create()
    {
        const int n = int(id.size());
        drv = g6d_create(n, id.data(), mass.data(),
                         reinterpret_cast<const double(*)[3]>(pos.data()),
                         reinterpret_cast<const double(*)[3]>(vel.data()), eps2, eta, dt_fac,
                         dt_limit, dt_median, precision, 1, gpu_id < 0 ? 0 : gpu_id);
        state_fresh = false;
        dirty = false;
        return drv ? 0 : -1;
    }
    // The driver rebuilt from the host copy at the system time.
    int recommit()
    {
        if (!drv) return -1;
        if (refresh()) return -1;
        long long b = 0, p = 0;
        double ts = 0.0;
        if (g6d_counters(drv, &b, &p, &ts) < 0) return -1;
        g6d_free(drv);
        drv = nullptr;
        blocks_offset += b;
        steps_offset += p;
        time_offset += ts;
        if (id.empty()) return 0;
        return create();
    }
    int index_of(int pid) const
    {
        auto it = slot.find(pid);
        return it == slot.end() ? -1 : it->second;
    }
    // device state -> host copies (raw state at each particle's own time)
    int refresh()
    {
        if (!drv || state_fresh) return 0;
        const size_t n = id.size();
        t.resize(n);
        dt.resize(n);
        if (g6d_state_raw(drv, reinterpret_cast<double(*)[3]>(pos.data()),
                          reinterpret_cast<double(*)[3]>(vel.data())) < 0 ||
            g6d_steps(drv, t.data(), dt.data()) < 0)
            return -1;
        state_fresh = true;
        return 0;
    }
};

Code code;

// One call of a function: typed access to its inputs and outputs.
struct Call {
    const Message &in;
    Message &out;
    int cc;      // call count
    int nio;     // int outputs including the result
    int ndo;     // double outputs
    int i = 0;   // current call
    int ii(int k) const { return in.ints[size_t(k) * cc + i]; }
    double di(int k) const { return in.doubles[size_t(k) * cc + i]; }
    int32_t &io(int k) { return out.ints[size_t(k + 1) * cc + i]; }  // result is int 0
    double &dout(int k) { return out.doubles[size_t(k) * cc + i]; }
};

struct Function {
    const char *name;
    int n_int_in, n_double_in, n_int_out, n_double_out;
    std::function<int(Call &)> body;
};

int not_committed() { return code.drv ? 0 : -1; }

std::vector<Function> functions()
{
    std::vector<Function> f;
    auto getd = [&](const char *name, double *p) {
        f.push_back({name, 0, 0, 0, 1, [p](Call &c) { c.dout(0) = *p; return 0; }});
    };
    auto setd = [&](const char *name, double *p) {
        f.push_back({name, 0, 1, 0, 0, [p](Call &c) { *p = c.di(0); return 0; }});
    };
    auto geti = [&](const char *name, int *p) {
        f.push_back({name, 0, 0, 1, 0, [p](Call &c) { c.io(0) = *p; return 0; }});
    };
    auto seti = [&](const char *name, int *p) {
        f.push_back({name, 1, 0, 0, 0, [p](Call &c) { *p = c.ii(0); return 0; }});
    };
    f.push_back({"initialize_code", 0, 0, 0, 0, [](Call &) { code.reset(); return 0; }});
    f.push_back({"cleanup_code", 0, 0, 0, 0, [](Call &) { code.reset(); return 0; }});
    f.push_back({"commit_parameters", 0, 0, 0, 0, [](Call &) { return 0; }});
    f.push_back({"recommit_parameters", 0, 0, 0, 0, [](Call &) { return 0; }});
    getd("get_eps2", &code.eps2);
    setd("set_eps2", &code.eps2);
    getd("get_eta", &code.eta);
    setd("set_eta", &code.eta);
    getd("get_begin_time", &code.begin_time);
    setd("set_begin_time", &code.begin_time);
    getd("get_initial_timestep_fac", &code.dt_fac);
    setd("set_initial_timestep_fac", &code.dt_fac);
    getd("get_initial_timestep_limit", &code.dt_limit);
    setd("set_initial_timestep_limit", &code.dt_limit);
    getd("get_initial_timestep_median", &code.dt_median);
    setd("set_initial_timestep_median", &code.dt_median);
    geti("get_gpu", &code.use_gpu);
    seti("set_gpu", &code.use_gpu);
    geti("get_gpu_id", &code.gpu_id);
    seti("set_gpu_id", &code.gpu_id);
    // force arithmetic of the block steps: 0 FP64, 1 double-single/FP32
    geti("get_force_precision", &code.precision);
    f.push_back({"set_force_precision", 1, 0, 0, 0, [](Call &c) {
                     const int p = c.ii(0);
                     if (p != G6X_FP64 && p != G6X_MIXED) return -1;
                     code.precision = p;
                     return code.drv ? (g6d_set_precision(code.drv, p) < 0 ? -1 : 0) : 0;
                 }});
    // new_particle(out index, mass, x, y, z, vx, vy, vz, radius, id): ids
    // from 1 up unless a free id >= 0 is asked for (jdata.cc:190-230)
    f.push_back({"new_particle", 1, 8, 1, 0, [](Call &c) {
                     if (code.drv) {  // the new body joins at the next recommit
                         if (code.refresh()) return -1;
                         code.dirty = true;
                     }
                     int pid = c.ii(0);
                     if (pid < 0 || code.slot.count(pid)) {
                         while (code.slot.count(code.next_id)) code.next_id++;
                         pid = code.next_id++;
                     }
                     code.slot[pid] = int(code.id.size());
                     if (code.drv) {  // time and step are set by the recommit
                         code.t.push_back(0.0);
                         code.dt.push_back(0.0);
                     }
                     code.id.push_back(pid);
                     code.mass.push_back(c.di(0));
                     for (int k = 0; k < 3; k++) code.pos.push_back(c.di(1 + k));
                     for (int k = 0; k < 3; k++) code.vel.push_back(c.di(4 + k));
                     code.radius.push_back(c.di(7));
                     c.io(0) = pid;
                     return 0;
                 }});
    f.push_back({"delete_particle", 1, 0, 0, 0, [](Call &c) {
                     const int j = code.index_of(c.ii(0));
                     if (j < 0) return -1;
                     if (code.drv) {
                         if (code.refresh()) return -1;
                         code.dirty = true;
                         code.t.erase(code.t.begin() + j);
                         code.dt.erase(code.dt.begin() + j);
                     }
                     code.id.erase(code.id.begin() + j);
                     code.mass.erase(code.mass.begin() + j);
                     code.radius.erase(code.radius.begin() + j);
                     code.pos.erase(code.pos.begin() + 3 * j, code.pos.begin() + 3 * j + 3);
                     code.vel.erase(code.vel.begin() + 3 * j, code.vel.begin() + 3 * j + 3);
                     code.reindex();
                     return 0;
                 }});
    // setters (interface.cc:466-580): ph4 writes its arrays; here the host
    // copy, which the next recommit loads (radius alone does not enter the
    // dynamics)
    auto edit = [&](const char *name, int nd, int what) {
        f.push_back({name, 1, nd, 0, 0, [what](Call &c) {
                         const int j = code.index_of(c.ii(0));
                         if (j < 0) return -1;
                         if (code.drv && code.refresh()) return -1;
                         int k = 0;
                         if (what == 0 || what == 1) code.mass[j] = c.di(k++);
                         if (what == 0 || what == 3)
                             for (int q = 0; q < 3; q++) code.pos[3 * j + q] = c.di(k++);
                         if (what == 0 || what == 4)
                             for (int q = 0; q < 3; q++) code.vel[3 * j + q] = c.di(k++);
                         if (what == 0 || what == 2) code.radius[j] = c.di(k++);
                         if (code.drv && what != 2) code.dirty = true;
                         return 0;
                     }});
    };
    edit("set_state", 8, 0);
    edit("set_mass", 1, 1);
    edit("set_radius", 1, 2);
    edit("set_position", 3, 3);
    edit("set_velocity", 3, 4);
    f.push_back({"commit_particles", 0, 0, 0, 0, [](Call &) {
                     if (code.drv || code.id.empty()) return -1;
                     code.time_offset = 0.0;
                     code.blocks_offset = code.steps_offset = 0;
                     if (code.create()) {
                         fprintf(stderr, "g6_worker: commit_particles: no GPU or a HIP error\n");
                         return -1;
                     }
                     return 0;
                 }});
    f.push_back({"recommit_particles", 0, 0, 0, 0, [](Call &) {
                     if (code.recommit()) {
                         fprintf(stderr, "g6_worker: recommit_particles: not committed, no GPU "
                                         "or a HIP error\n");
                         return -1;
                     }
                     return 0;
                 }});
    f.push_back({"synchronize_model", 0, 0, 0, 0, [](Call &) {
                     if (!code.drv) return 0;
                     if (code.dirty && code.recommit()) return -1;
                     code.state_fresh = false;
                     return g6d_synchronize(code.drv) < 0 ? -1 : 0;
                 }});
    f.push_back({"evolve_model", 0, 1, 0, 0, [](Call &c) {
                     if (not_committed()) return -1;
                     if (code.dirty && code.recommit()) return -1;
                     int reached = 0;
                     const long long r =
                         g6d_run(code.drv, c.di(0) - code.begin_time - code.time_offset, 0, 64,
                                 &reached);
                     code.state_fresh = false;
                     if (r < 0) {
                         fprintf(stderr, "g6_worker: evolve_model: %s\n",
                                 g6d_last_error(code.drv));
                         return -1;
                     }
                     return 0;
                 }});
    f.push_back({"get_time", 0, 0, 0, 1, [](Call &c) {
                     long long b = 0, p = 0;
                     double ts = 0.0;
                     if (code.counters(b, p, ts)) return -1;
                     c.dout(0) = ts + code.begin_time;
                     return 0;
                 }});
    auto counter = [&](const char *name, bool blocks) {
        f.push_back({name, 0, 0, 1, 0, [blocks](Call &c) {
                         long long b = 0, p = 0;
                         double ts = 0.0;
                         if (code.counters(b, p, ts)) return -1;
                         c.io(0) = int(blocks ? b : p);
                         return 0;
                     }});
    };
    counter("get_block_steps", true);
    counter("get_total_steps", false);
    f.push_back({"get_number_of_particles", 0, 0, 1, 0, [](Call &c) {
                     c.io(0) = int(code.id.size());
                     return 0;
                 }});
    f.push_back({"get_total_mass", 0, 0, 0, 1, [](Call &c) {
                     double m = 0.0;
                     for (double q : code.mass) m += q;
                     c.dout(0) = m;
                     return 0;
                 }});
    auto energy = [&](const char *name, bool kinetic) {
        f.push_back({name, 0, 0, 0, 1, [kinetic](Call &c) {
                         if (not_committed()) return -1;
                         if (code.dirty && code.recommit()) return -1;
                         double k = 0.0, u = 0.0;
                         if (g6d_energy(code.drv, &k, &u) < 0) return -1;
                         c.dout(0) = kinetic ? k : u;
                         return 0;
                     }});
    };
    energy("get_kinetic_energy", true);
    energy("get_potential_energy", false);
    // per-particle getters (arrays): index in, values out
    f.push_back({"get_state", 1, 0, 0, 8, [](Call &c) {
                     if (code.refresh()) return -1;
                     const int j = code.index_of(c.ii(0));
                     if (j < 0) return -1;
                     c.dout(0) = code.mass[j];
                     for (int k = 0; k < 3; k++) c.dout(1 + k) = code.pos[3 * j + k];
                     for (int k = 0; k < 3; k++) c.dout(4 + k) = code.vel[3 * j + k];
                     c.dout(7) = code.radius[j];
                     return 0;
                 }});
    f.push_back({"get_mass", 1, 0, 0, 1, [](Call &c) {
                     const int j = code.index_of(c.ii(0));
                     if (j < 0) return -1;
                     c.dout(0) = code.mass[j];
                     return 0;
                 }});
    f.push_back({"get_radius", 1, 0, 0, 1, [](Call &c) {
                     const int j = code.index_of(c.ii(0));
                     if (j < 0) return -1;
                     c.dout(0) = code.radius[j];
                     return 0;
                 }});
    auto vec3 = [&](const char *name, bool position) {
        f.push_back({name, 1, 0, 0, 3, [position](Call &c) {
                         if (code.refresh()) return -1;
                         const int j = code.index_of(c.ii(0));
                         if (j < 0) return -1;
                         const std::vector<double> &a = position ? code.pos : code.vel;
                         for (int k = 0; k < 3; k++) c.dout(k) = a[3 * j + k];
                         return 0;
                     }});
    };
    vec3("get_position", true);
    vec3("get_velocity", false);
    f.push_back({"get_particle_timestep", 1, 0, 0, 1, [](Call &c) {
                     if (not_committed()) return -1;
                     if (code.dirty && code.recommit()) return -1;
                     if (code.refresh()) return -1;
                     const int j = code.index_of(c.ii(0));
                     if (j < 0) return -1;
                     c.dout(0) = code.dt[j];
                     return 0;
                 }});
    return f;
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
void handle(const std::map<int, Function> &table, const Message &in, Message &out)
{
    const int cc = in.header[H_CALL_COUNT];
    out = Message();
    std::memcpy(out.header, in.header, sizeof(out.header));
    out.header[H_FLAGS] = 0;
    auto fail = [&](const std::string &why) {
        out.header[H_FLAGS] = kErrorFlag;
        out.ints.clear();
        out.doubles.clear();
        out.strings = {why};
    };
    auto it = table.find(in.header[H_FUNCTION_ID]);
    if (it == table.end()) {
        fail("unknown function id: " + std::to_string(in.header[H_FUNCTION_ID]) + "\n");
        return;
    }
    const Function &f = it->second;
    if (cc < 1 || in.ints.size() != size_t(f.n_int_in) * cc ||
        in.doubles.size() != size_t(f.n_double_in) * cc) {
        fail(std::string("bad argument counts for ") + f.name);
        return;
    }
    out.ints.assign(size_t(1 + f.n_int_out) * cc, 0);
    out.doubles.assign(size_t(f.n_double_out) * cc, 0.0);
    Call c{in, out, cc, 1 + f.n_int_out, f.n_double_out};
    for (c.i = 0; c.i < cc; c.i++) out.ints[size_t(c.i)] = f.body(c);
}

int connect_to(int port, const char *host)
{
    addrinfo hints{}, *res = nullptr;
    hints.ai_family = AF_INET;
    hints.ai_socktype = SOCK_STREAM;
    const std::string p = std::to_string(port);
    if (getaddrinfo(host, p.c_str(), &hints, &res) != 0 || !res) return -1;
    const int fd = socket(AF_INET, SOCK_STREAM, 0);
    if (fd < 0 || connect(fd, res->ai_addr, res->ai_addrlen) != 0) {
        freeaddrinfo(res);
        if (fd >= 0) close(fd);
        return -1;
    }
    freeaddrinfo(res);
    int on = 1;
    setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &on, sizeof(on));
    return fd;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 4 || std::strcmp(argv[3], "false") != 0) {
        fprintf(stderr, "usage: %s PORT HOST false   (socket channel, no MPI)\n", argv[0]);
        return 2;
    }
    const int fd = connect_to(std::atoi(argv[1]), argv[2]);
    if (fd < 0) {
        fprintf(stderr, "g6_worker: cannot connect to %s:%s\n", argv[2], argv[1]);
        return 1;
    }
    std::map<int, Function> table;
    for (Function &f : functions()) table[function_id(f.name)] = f;
    Message in, out;
    for (;;) {
        if (!receive(fd, in)) break;
        const bool stop = in.header[H_FUNCTION_ID] == 0;
        if (stop) {
            out = Message();
            std::memcpy(out.header, in.header, sizeof(out.header));
            out.header[H_FLAGS] = 0;
        } else {
            handle(table, in, out);
        }
        if (!reply(fd, out) || stop) break;
    }
    code.reset();
    close(fd);
    return 0;
}
