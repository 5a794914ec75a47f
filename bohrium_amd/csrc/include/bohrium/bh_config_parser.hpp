/* INI configuration with BH_<SECTION>_<OPTION> environment overrides.
 *
 * Lookup rules follow core/bh_config_parser.cpp:44-172 of the reference:
 * BH_CONFIG, then ~/.bohrium/config.ini, then the system paths; the stack is the
 * comma/space separated list under [stacks] named by BH_STACK ("default" when unset);
 * an option is read from the environment first, then from the file, with one
 * pair of surrounding quotes removed.  This implementation adds one more
 * fallback at the end of the search: the config.ini shipped next to the
 * package's libraries (passed in by the bridge), since this build is not installed
 * into /etc.
 */
#pragma once

#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace bohrium {

class ConfigError : public std::runtime_error {
public:
    explicit ConfigError(const std::string &msg) : std::runtime_error(msg) {}
};

class ConfigParser {
    std::map<std::string, std::map<std::string, std::string> > _ini;
    std::vector<std::string> _stack_list;
    std::string _default_section;

    std::string lookup(const std::string &section, const std::string &option) const;

public:
    std::string file_path;
    std::string file_dir;
    const int stack_level;

    explicit ConfigParser(int stack_level);

    /* Set the fallback config file used when no other config is found. */
    static void setFallbackPath(const std::string &path);

    bool has(const std::string &section, const std::string &option) const;

    template<typename T>
    T get(const std::string &section, const std::string &option) const {
        std::string s = lookup(section, option);
        return convert<T>(s, section, option);
    }
    template<typename T>
    T get(const std::string &option) const { return get<T>(_default_section, option); }

    template<typename T>
    T defaultGet(const std::string &option, const T &default_value) const {
        if (!has(_default_section, option)) return default_value;
        return get<T>(_default_section, option);
    }

    std::vector<std::string> getList(const std::string &section, const std::string &option) const;
    std::vector<std::string> getList(const std::string &option) const { return getList(_default_section, option); }
    std::vector<std::string> defaultGetList(const std::string &option, const std::vector<std::string> &def) const {
        if (!has(_default_section, option)) return def;
        return getList(_default_section, option);
    }

    std::string getChildLibraryPath() const;
    const std::string &getName() const { return _default_section; }
    const std::vector<std::string> &getStackList() const { return _stack_list; }

private:
    template<typename T>
    static T convert(const std::string &s, const std::string &section, const std::string &option) {
        std::istringstream is(s);
        T ret;
        if (!(is >> ret)) throw ConfigError("config: cannot parse " + section + "." + option + " = '" + s + "'");
        return ret;
    }
};

template<>
inline std::string ConfigParser::convert<std::string>(const std::string &s, const std::string &, const std::string &) {
    return s;
}

template<>
inline bool ConfigParser::convert<bool>(const std::string &s, const std::string &section, const std::string &option) {
    if (s == "true" || s == "True" || s == "1" || s == "yes" || s == "on") return true;
    if (s == "false" || s == "False" || s == "0" || s == "no" || s == "off") return false;
    throw ConfigError("config: cannot parse bool " + section + "." + option + " = '" + s + "'");
}

} // namespace bohrium
