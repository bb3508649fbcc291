// config.cpp -- reader of the reference's configuration files (rwnmr.config,
// cpmg.config, pfgse.config, multitau.config, uct.config, ga.config): a
// key/value store filled file by file, with the reference's line grammar
// and value conversions.
//
// Reference behaviour restated (paths under src/backend/ of RandomWalk_NMR):
//   line grammar: every ": " of a line separates a key candidate from the
//   rest of the line, which is its value (no trimming)
//                                   B/ConfigFiles/rwnmr_config.cpp:79-145 (same in every *_config.cpp)
//   defaults file first, then the user's file
//                                   B/ConfigFiles/rwnmr_config.cpp:15-23
//   numbers std::stod / std::stoi, booleans == "true"
//                                   B/ConfigFiles/rwnmr_config.cpp:165-369
//   lists "{a, b, ...}" split at "," and std::stod per item; RHO appends
//   to the values read so far (every file adds)
//                                   B/ConfigFiles/rwnmr_config.cpp:178-202
#include <stdint.h>
#include <string.h>

#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/rwnmr/rwnmr_config.h"

struct rwnmr_config_file {
    std::map<std::string, std::vector<std::string>> values;   // every occurrence, in reading order
};

namespace {

bool to_double(const std::string &s, double &v)
{
    try {
        v = std::stod(s);
        return true;
    } catch (...) {
        return false;
    }
}

// "{a, b, c}" -> a, b, c; anything else -> one value
bool parse_list(const std::string &s, std::vector<double> &out)
{
    if (s.size() >= 2 && s.front() == '{' && s.back() == '}') {
        std::string body = s.substr(1, s.size() - 2);
        size_t pos;
        double v;
        while ((pos = body.find(',')) != std::string::npos) {
            if (!to_double(body.substr(0, pos), v))
                return false;
            out.push_back(v);
            body.erase(0, pos + 1);
        }
        if (!to_double(body, v))
            return false;
        out.push_back(v);
        return true;
    }
    double v;
    if (!to_double(s, v))
        return false;
    out.push_back(v);
    return true;
}

const std::string *last_value(const rwnmr_config_file *c, const char *key)
{
    if (!c || !key)
        return nullptr;
    auto it = c->values.find(key);
    if (it == c->values.end() || it->second.empty())
        return nullptr;
    return &it->second.back();
}

}  // namespace

extern "C" {

rwnmr_status rwnmr_config_create(rwnmr_config_file **out)
{
    if (!out)
        return RWNMR_ERR_INVALID;
    *out = new (std::nothrow) rwnmr_config_file();
    return *out ? RWNMR_OK : RWNMR_ERR_ALLOC;
}

void rwnmr_config_destroy(rwnmr_config_file *c) { delete c; }

rwnmr_status rwnmr_config_read(rwnmr_config_file *c, const char *path)
{
    if (!c || !path)
        return RWNMR_ERR_INVALID;
    std::ifstream f(path);
    if (!f)
        return RWNMR_ERR_INVALID;
    std::string line;
    while (std::getline(f, line)) {
        std::string s = line;
        size_t pos;
        while ((pos = s.find(": ")) != std::string::npos) {
            c->values[s.substr(0, pos)].push_back(s.substr(pos + 2));
            s.erase(0, pos + 2);
        }
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_config_set(rwnmr_config_file *c, const char *key, const char *value)
{
    if (!c || !key || !value)
        return RWNMR_ERR_INVALID;
    c->values[key].push_back(value);
    return RWNMR_OK;
}

int32_t rwnmr_config_has(const rwnmr_config_file *c, const char *key) { return last_value(c, key) != nullptr; }

rwnmr_status rwnmr_config_string(const rwnmr_config_file *c, const char *key, char *buf, uint32_t buflen)
{
    const std::string *v = last_value(c, key);
    if (!v || !buf || v->size() + 1 > buflen)
        return RWNMR_ERR_INVALID;
    memcpy(buf, v->c_str(), v->size() + 1);
    return RWNMR_OK;
}

rwnmr_status rwnmr_config_double(const rwnmr_config_file *c, const char *key, double *value)
{
    const std::string *v = last_value(c, key);
    if (!v || !value)
        return RWNMR_ERR_INVALID;
    return to_double(*v, *value) ? RWNMR_OK : RWNMR_ERR_INVALID;
}

rwnmr_status rwnmr_config_int(const rwnmr_config_file *c, const char *key, int64_t *value)
{
    const std::string *v = last_value(c, key);
    if (!v || !value)
        return RWNMR_ERR_INVALID;
    try {
        *value = (int64_t)std::stoll(*v);
    } catch (...) {
        return RWNMR_ERR_INVALID;
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_config_bool(const rwnmr_config_file *c, const char *key, int32_t *value)
{
    const std::string *v = last_value(c, key);
    if (!v || !value)
        return RWNMR_ERR_INVALID;
    *value = *v == "true";
    return RWNMR_OK;
}

rwnmr_status rwnmr_config_list(const rwnmr_config_file *c, const char *key, int32_t append, double *out,
                               uint32_t capacity, uint32_t *n)
{
    if (!c || !key || !n)
        return RWNMR_ERR_INVALID;
    auto it = c->values.find(key);
    if (it == c->values.end() || it->second.empty())
        return RWNMR_ERR_INVALID;
    std::vector<double> vals;
    if (append) {
        for (const std::string &s : it->second)
            if (!parse_list(s, vals))
                return RWNMR_ERR_INVALID;
    } else if (!parse_list(it->second.back(), vals)) {
        return RWNMR_ERR_INVALID;
    }
    *n = (uint32_t)vals.size();
    if (out) {
        if (vals.size() > capacity)
            return RWNMR_ERR_INVALID;
        memcpy(out, vals.data(), vals.size() * sizeof(double));
    }
    return RWNMR_OK;
}

}  // extern "C"
