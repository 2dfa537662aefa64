// Persistent autotuner cache (tune_cache.h).  The file is JSON written one
// entry per line by tune_store,
//   {
//   "format": "lassie-tune-1",
//   "entries": {
//   "<key>": {"choice": 2, "trial_ms": [12.5, 11.9, 12.1]},
//   ...
//   }
//   }
// and read back line by line; keys hold no quotes or backslashes.
#include "tune_cache.h"

#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace lassie {

namespace {

std::mutex g_mu;
const char *kPrefix = "\": {\"choice\": ";

// key -> the entry's text after the key, as written ({"choice": ...})
std::map<std::string, std::string> read_entries(const std::string &path)
{
    std::map<std::string, std::string> out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.size() < 4 || line[0] != '"') continue;
        const size_t q = line.find(kPrefix);
        if (q == std::string::npos) continue;
        std::string rest = line.substr(q + 3);
        if (!rest.empty() && rest.back() == ',') rest.pop_back();
        out[line.substr(1, q - 1)] = rest;
    }
    return out;
}

void make_dirs(const std::string &path)
{
    for (size_t i = 1; i < path.size(); i++)
        if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
}

}  // namespace

std::string tune_cache_path()
{
    const char *p = std::getenv("LASSIE_TUNE_CACHE");
    if (p && *p) return std::string(p) == "off" ? std::string() : std::string(p);
    const char *x = std::getenv("XDG_CACHE_HOME");
    if (x && *x) return std::string(x) + "/lassie/tune.json";
    const char *h = std::getenv("HOME");
    if (h && *h) return std::string(h) + "/.cache/lassie/tune.json";
    return std::string();
}

bool tune_lookup(const std::string &key, int *choice)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    const auto entries = read_entries(path);
    const auto it = entries.find(key);
    if (it == entries.end()) return false;
    const size_t c = it->second.find("\"choice\": ");
    if (c == std::string::npos) return false;
    char *end = nullptr;
    const long v = std::strtol(it->second.c_str() + c + 10, &end, 10);
    if (end == it->second.c_str() + c + 10) return false;
    *choice = (int)v;
    return true;
}

void tune_store(const std::string &key, int choice, const float *trial_ms, int n_trials)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto entries = read_entries(path);
    std::ostringstream e;
    e << "{\"choice\": " << choice << ", \"trial_ms\": [";
    for (int i = 0; i < n_trials; i++) e << (i ? ", " : "") << trial_ms[i];
    e << "]}";
    entries[key] = e.str();
    make_dirs(path);
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) return;
        f << "{\n\"format\": \"lassie-tune-1\",\n\"entries\": {\n";
        size_t i = 0;
        for (const auto &kv : entries) f << '"' << kv.first << "\": " << kv.second << (++i < entries.size() ? ",\n" : "\n");
        f << "}\n}\n";
        if (!f) return;
    }
    (void)std::rename(tmp.c_str(), path.c_str());
}

}  // namespace lassie
