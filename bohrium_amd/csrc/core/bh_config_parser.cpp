#include <bohrium/bh_config_parser.hpp>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sys/stat.h>

namespace bohrium {

namespace {
std::string g_fallback_path;

bool file_exists(const std::string &p) {
    struct stat st;
    return !p.empty() && stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string trim(const std::string &s) {
    size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    size_t e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

// Search order of the reference (core/bh_config_parser.cpp:44-120), then the packaged file.
std::string find_config_path() {
    const char *env = std::getenv("BH_CONFIG");
    if (env != nullptr && file_exists(env)) return env;
    const char *home = std::getenv("HOME");
    if (home != nullptr) {
        std::string p = std::string(home) + "/.bohrium/config.ini";
        if (file_exists(p)) return p;
    }
    for (const char *p : {"/usr/local/etc/bohrium/config.ini", "/usr/etc/bohrium/config.ini",
                          "/etc/bohrium/config.ini"}) {
        if (file_exists(p)) return p;
    }
    if (file_exists(g_fallback_path)) return g_fallback_path;
    throw ConfigError("Bohrium could not find the config file (BH_CONFIG, ~/.bohrium/config.ini, /etc/bohrium)");
}

std::string env_name(const std::string &section, const std::string &option) {
    std::string s = "BH_" + section + "_" + option;
    for (char &c : s) {
        c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
        if (c == '-' || c == ' ') c = '_';
    }
    return s;
}
} // namespace

void ConfigParser::setFallbackPath(const std::string &path) { g_fallback_path = path; }

ConfigParser::ConfigParser(int stack_level) : stack_level(stack_level) {
    file_path = find_config_path();
    const size_t slash = file_path.find_last_of('/');
    file_dir = slash == std::string::npos ? "." : file_path.substr(0, slash);

    std::ifstream in(file_path);
    std::string line, section;
    while (std::getline(in, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#' || t[0] == ';') continue;
        if (t.front() == '[' && t.back() == ']') {
            section = trim(t.substr(1, t.size() - 2));
            continue;
        }
        const size_t eq = t.find('=');
        if (eq == std::string::npos) continue;
        _ini[section][trim(t.substr(0, eq))] = trim(t.substr(eq + 1));
    }

    const char *env = std::getenv("BH_STACK");
    const std::string stack_name = env == nullptr ? "default" : env;
    _stack_list = getList("stacks", stack_name);
    if (stack_level >= static_cast<int>(_stack_list.size()) || stack_level < -1)
        throw ConfigError("ConfigParser: stack level is out of bound");
    _default_section = stack_level == -1 ? "bridge" : _stack_list[stack_level];
}

bool ConfigParser::has(const std::string &section, const std::string &option) const {
    if (std::getenv(env_name(section, option).c_str()) != nullptr) return true;
    auto s = _ini.find(section);
    return s != _ini.end() && s->second.count(option) > 0;
}

std::string ConfigParser::lookup(const std::string &section, const std::string &option) const {
    const char *env = std::getenv(env_name(section, option).c_str());
    if (env != nullptr && *env != '\0') return env;
    auto s = _ini.find(section);
    if (s == _ini.end() || s->second.count(option) == 0)
        throw ConfigError("config: no option " + section + "." + option + " in " + file_path);
    std::string ret = s->second.at(option);
    if (ret.size() >= 2 && (ret.front() == '"' || ret.front() == '\'') && (ret.back() == '"' || ret.back() == '\''))
        return ret.substr(1, ret.size() - 2);
    return ret;
}

std::vector<std::string> ConfigParser::getList(const std::string &section, const std::string &option) const {
    std::vector<std::string> ret;
    std::string cur;
    for (char c : lookup(section, option)) {
        if (c == ',' || c == ' ' || c == '\t') {
            if (!cur.empty()) ret.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) ret.push_back(cur);
    return ret;
}

std::string ConfigParser::getChildLibraryPath() const {
    if (static_cast<int>(_stack_list.size()) <= stack_level + 1) return std::string();
    std::string p = get<std::string>(_stack_list[stack_level + 1], "impl");
    if (!p.empty() && p[0] != '/') p = file_dir + "/" + p;
    return p;
}

} // namespace bohrium
