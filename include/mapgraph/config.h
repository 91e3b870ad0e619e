// config.h -- typed parameter registry with the reference's parameter names.
//
// Mirrors Config (Algorithms/BFS/config.h:70-175, config.cpp:30-115): parameters are
// registered with a default, set from "k=v k=v" strings (-p) or a file of such
// lines (-c), and read with getParameter<T>(name).  Differences, by design:
//   * an unknown name is an error (reference: dereferences end(), config.cpp:88-94);
//   * "true"/"false" parse as such (reference getValue<bool> is inverted,
//     getvalue.h:35-45);
//   * the registry is per object, not a process-wide static.
// Default parameters are those of Algorithms/*/register.h.
#pragma once

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace mapgraph {

class Config {
  public:
    struct Param {
        std::string value;
        std::string description;
    };

    Config() { register_defaults(); }

    void registerParameter(const std::string &name, const std::string &description, const std::string &value) {
        params_[name] = Param{value, description};
    }
    template <class T>
    void registerParameter(const std::string &name, const std::string &description, T value) {
        std::ostringstream os;
        os << value;
        registerParameter(name, description, os.str());
    }

    bool has(const std::string &name) const { return params_.count(name) != 0; }

    // "name=value" (or "name=value name2=value2 ...", separated by spaces or commas)
    void parseParameterString(const std::string &s) {
        std::string tok;
        for (size_t i = 0; i <= s.size(); ++i) {
            const char ch = i < s.size() ? s[i] : ' ';
            if (ch == ' ' || ch == ',' || ch == '\t' || ch == '\n' || ch == '\r') {
                if (!tok.empty()) setParameter(tok);
                tok.clear();
            } else {
                tok.push_back(ch);
            }
        }
    }

    void parseFile(const std::string &path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open config file '" + path + "'");
        std::string line;
        while (std::getline(in, line)) {
            const size_t hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            parseParameterString(line);
        }
    }

    void setParameter(const std::string &assignment) {
        const size_t eq = assignment.find('=');
        if (eq == std::string::npos) throw std::invalid_argument("parameter '" + assignment + "' is not name=value");
        setParameter(assignment.substr(0, eq), assignment.substr(eq + 1));
    }

    void setParameter(const std::string &name, const std::string &value) {
        auto it = params_.find(name);
        if (it == params_.end()) throw std::invalid_argument("parameter '" + name + "' is not registered");
        it->second.value = value;
    }

    template <class T>
    T getParameter(const std::string &name) const {
        auto it = params_.find(name);
        if (it == params_.end()) throw std::invalid_argument("parameter '" + name + "' is not registered");
        return parse<T>(it->second.value);
    }

    void printOptions(FILE *f = stdout) const {
        for (const auto &p : params_) fprintf(f, "           %s: %s\n", p.first.c_str(), p.second.description.c_str());
    }

  private:
    template <class T>
    static T parse(const std::string &v);

    void register_defaults() {
        registerParameter<int>("src", "Starting vertex (default 1)", 1);
        registerParameter<int>("verbose", "Print out frontier size in each iteration (default 0)", 0);
        registerParameter<int>("origin", "The origin (0 or 1) for the starting vertices (default 1)", 1);
        registerParameter<int>("directed", "The graph is directed (default 1)", 1);
        registerParameter<int>("device", "The device to use (default 0)", 0);
        registerParameter<int>("iter_num", "The number of iterations to perform (default INT_MAX)", INT_MAX);
        registerParameter<int>("num_src", "The number of starting vertices when random sources is specified (default 1)", 1);
        registerParameter<int>("run_CPU", "Run CPU implementation for testing (default 0)", 0);
        registerParameter<int>("with_value", "Whether to load edge values from market file (default 0)", 0);
        registerParameter<double>("max_queue_sizing", "The frontier queue size is this value times the number of vertices in the graph (default 1.5)", 1.5);
        registerParameter<int>("threshold", "When frontier size is larger than threshold, two-phase strategy is used otherwise dynamic scheduling it used (default 10000)", 10000);
        registerParameter<int>("stats", "Print out degree histogram (default 0)", 0);
        registerParameter<double>("tol", "The tolerance for correctness check (default 0.1)", 0.1);
    }

    std::map<std::string, Param> params_;
};

template <>
inline int Config::parse<int>(const std::string &v) {
    return (int)std::strtol(v.c_str(), nullptr, 10);
}
template <>
inline double Config::parse<double>(const std::string &v) {
    return std::strtod(v.c_str(), nullptr);
}
template <>
inline float Config::parse<float>(const std::string &v) {
    return std::strtof(v.c_str(), nullptr);
}
template <>
inline bool Config::parse<bool>(const std::string &v) {
    return v == "true" || v == "1";
}
template <>
inline std::string Config::parse<std::string>(const std::string &v) {
    return v;
}

}  // namespace mapgraph
