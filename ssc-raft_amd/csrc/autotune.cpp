// autotune.cpp -- the library's tile autotuner.
//
// A tuning key names a kernel family and the geometry that decides its
// speed ("bp/n=2048/A=2048/W=2912/store=0/mode=0", "proj/..."); its value is
// the chosen variant (an angle chunk of the backprojector, a ray-chunk count
// of the staged projector).  The first call that needs a key times every
// candidate on the call's own buffers with HIP events on the call's stream
// (two runs each, the faster counts) and keeps the fastest.  Choices live in
// memory for the process and, when RAFT_AUTOTUNE_CACHE names a file, in that
// JSON file across processes:
//     {"<key>": {"choice": 32, "ms": [[16, 0.91], [32, 0.87], [48, 0.95]]}, ...}
// RAFT_AUTOTUNE=0 turns tuning off (every key takes its default, the first
// candidate).  Every variant a key chooses between computes the same values
// bit for bit, so a choice changes speed only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

#include "raft_internal.hpp"

namespace raft {

namespace {

struct TuneRec {
    int choice = 0;
    std::vector<std::pair<int, float>> ms;
};

std::mutex g_tune_mu;
std::map<std::string, TuneRec> g_tune;
bool g_tune_loaded = false;

const char* cache_path()
{
    const char* p = std::getenv("RAFT_AUTOTUNE_CACHE");
    return (p && p[0]) ? p : nullptr;
}

// minimal reader of the file format above: "key": {"choice": N ...}
void load_locked()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* path = cache_path();
    if (!path) return;
    FILE* f = std::fopen(path, "rb");
    if (!f) return;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    size_t pos = 0;
    for (;;) {
        const size_t k0 = text.find('"', pos);
        if (k0 == std::string::npos) break;
        const size_t k1 = text.find('"', k0 + 1);
        if (k1 == std::string::npos) break;
        const std::string key = text.substr(k0 + 1, k1 - k0 - 1);
        const size_t c = text.find("\"choice\"", k1);
        const size_t next = text.find("},", k1);
        if (c == std::string::npos) break;
        if (next != std::string::npos && c > next) { pos = next + 2; continue; }
        const size_t colon = text.find(':', c);
        if (colon == std::string::npos) break;
        const int v = std::atoi(text.c_str() + colon + 1);
        if (key.find('/') != std::string::npos && g_tune.find(key) == g_tune.end()) g_tune[key].choice = v;
        pos = (next == std::string::npos) ? text.size() : next + 2;
    }
}

std::string dump_locked()
{
    std::ostringstream o;
    o << "{";
    bool first = true;
    for (const auto& kv : g_tune) {
        o << (first ? "\n" : ",\n") << " \"" << kv.first << "\": {\"choice\": " << kv.second.choice << ", \"ms\": [";
        for (size_t i = 0; i < kv.second.ms.size(); ++i)
            o << (i ? ", " : "") << "[" << kv.second.ms[i].first << ", " << kv.second.ms[i].second << "]";
        o << "]}";
        first = false;
    }
    o << (first ? "}" : "\n}") << "\n";
    return o.str();
}

void save_locked()
{
    const char* path = cache_path();
    if (!path) return;
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const std::string s = dump_locked();
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    std::fclose(f);
    if (ok) std::rename(tmp.c_str(), path);
    else std::remove(tmp.c_str());
}

} // namespace

bool tune_enabled()
{
    const char* e = std::getenv("RAFT_AUTOTUNE");
    return !(e && e[0] == '0');
}

int tune_pick(const std::string& key, const std::vector<int>& cands, hipStream_t st,
              const std::function<void(int)>& launch)
{
    if (cands.empty()) fail(RAFT_EINVAL, "internal: no tuning candidates");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        load_locked();
        auto it = g_tune.find(key);
        if (it != g_tune.end())
            for (int c : cands)
                if (c == it->second.choice) return c;
    }
    if (!tune_enabled() || cands.size() == 1) return cands[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RAFT_HIP(hipEventCreate(&e0));
    RAFT_HIP(hipEventCreate(&e1));
    TuneRec rec;
    float best = 1e30f;
    try {
        for (int c : cands) {
            float tmin = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                RAFT_HIP(hipEventRecord(e0, st));
                launch(c);
                RAFT_HIP(hipEventRecord(e1, st));
                RAFT_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                RAFT_HIP(hipEventElapsedTime(&ms, e0, e1));
                tmin = std::min(tmin, ms);
            }
            rec.ms.push_back({c, tmin});
            if (tmin < best) {
                best = tmin;
                rec.choice = c;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = rec;
    save_locked();
    return rec.choice;
}

int tune_cached(const std::string& key, int dflt)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_locked();
    auto it = g_tune.find(key);
    return it == g_tune.end() ? dflt : it->second.choice;
}

} // namespace raft

extern "C" {

int raft_autotune_dump(char* buf, int len)
{
    std::lock_guard<std::mutex> lk(raft::g_tune_mu);
    raft::load_locked();
    const std::string s = raft::dump_locked();
    if (buf && len > 0) {
        const size_t n = std::min((size_t)(len - 1), s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return (int)s.size() + 1;
}

void raft_autotune_clear(void)
{
    std::lock_guard<std::mutex> lk(raft::g_tune_mu);
    raft::g_tune.clear();
    raft::g_tune_loaded = false;
}

} // extern "C"
