// Opm::ParameterGroup: the key=value parameter store generated Equelle programs
// construct from argv (PrintCUDABackendASTVisitor.cpp:45).  Accepts `key=value`
// arguments and parameter files (one `key=value` per line, '#' comments), later
// entries overriding earlier ones.
#ifndef EQUELLE_AMD_OPM_PARAMETERGROUP_HPP
#define EQUELLE_AMD_OPM_PARAMETERGROUP_HPP
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace Opm {

class ParameterGroup {
public:
    ParameterGroup() = default;
    ParameterGroup(int argc, char** argv, bool verify_syntax = true, bool enable_output = true) {
        (void)verify_syntax;
        (void)enable_output;
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a.find('=') != std::string::npos) insert(a);
            else readFile(a);
        }
    }
    bool has(const std::string& name) const { return kv_.count(name) != 0; }
    template <class T>
    T get(const std::string& name) const {
        auto it = kv_.find(name);
        if (it == kv_.end()) throw std::runtime_error("ParameterGroup: parameter " + name + " not found");
        return convert<T>(it->second);
    }
    template <class T>
    T getDefault(const std::string& name, const T& def) const {
        auto it = kv_.find(name);
        return it == kv_.end() ? def : convert<T>(it->second);
    }
    std::string getDefault(const std::string& name, const char* def) const {
        return getDefault<std::string>(name, std::string(def));
    }
    void insertParameter(const std::string& name, const std::string& value) { kv_[name] = value; }
    // All entries as "key=value key=value" (used to configure the device runtime).
    std::string toString() const {
        std::string s;
        for (const auto& e : kv_) s += e.first + "=" + e.second + " ";
        return s;
    }
    void readFile(const std::string& path) {
        std::ifstream is(path.c_str());
        if (!is) throw std::runtime_error("ParameterGroup: cannot read parameter file " + path);
        std::string line;
        while (std::getline(is, line)) {
            const size_t h = line.find('#');
            if (h != std::string::npos) line.erase(h);
            insert(line);
        }
    }

private:
    void insert(const std::string& a) {
        const size_t e = a.find('=');
        if (e == std::string::npos) return;
        std::string k = trim(a.substr(0, e)), v = trim(a.substr(e + 1));
        if (!k.empty()) kv_[k] = v;
    }
    static std::string trim(const std::string& s) {
        const size_t b = s.find_first_not_of(" \t\r\n");
        if (b == std::string::npos) return "";
        const size_t e = s.find_last_not_of(" \t\r\n");
        return s.substr(b, e - b + 1);
    }
    template <class T>
    static T convert(const std::string& s) {
        std::istringstream is(s);
        T v{};
        is >> v;
        return v;
    }
    std::map<std::string, std::string> kv_;
};

template <>
inline std::string ParameterGroup::convert<std::string>(const std::string& s) {
    return s;
}
template <>
inline bool ParameterGroup::convert<bool>(const std::string& s) {
    return s == "true" || s == "1" || s == "yes";
}

}  // namespace Opm
#endif
