// cli.h -- minimal "--name value" flag parser for the HRNN drivers.
// Flags, defaults and "required" markers mirror the reference drivers'
// program_options declarations; unknown flags are an error.
#pragma once

#include <cstdlib>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

class Cli {
    struct Opt {
        std::string value, help;
        bool required, given;
    };
    std::map<std::string, Opt> opts_;
    std::vector<std::string> order_;
    std::string title_;

 public:
    explicit Cli(std::string title) : title_(std::move(title)) {}
    void add(const std::string &name, const std::string &def, const std::string &help, bool required = false) {
        opts_[name] = Opt{def, help, required, false};
        order_.push_back(name);
    }
    // returns false when --help was asked
    bool parse(int argc, char **argv) {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "--help" || a == "-h") {
                usage();
                return false;
            }
            if (a.rfind("--", 0) != 0) throw std::runtime_error("unexpected argument: " + a);
            std::string name = a.substr(2), val;
            auto eq = name.find('=');
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
            } else {
                if (i + 1 >= argc) throw std::runtime_error("missing value for --" + name);
                val = argv[++i];
            }
            auto it = opts_.find(name);
            if (it == opts_.end()) throw std::runtime_error("unknown option --" + name);
            it->second.value = val;
            it->second.given = true;
        }
        for (auto &kv : opts_)
            if (kv.second.required && !kv.second.given)
                throw std::runtime_error("the option '--" + kv.first + "' is required but missing");
        return true;
    }
    void usage() const {
        std::cout << title_ << ":\n";
        for (auto &n : order_) {
            const Opt &o = opts_.at(n);
            std::cout << "  --" << n << (o.required ? " (required)" : " [=" + o.value + "]") << "  "
                      << o.help << "\n";
        }
    }
    std::string str(const std::string &n) const { return opts_.at(n).value; }
    int i(const std::string &n) const { return std::atoi(opts_.at(n).value.c_str()); }
    double f(const std::string &n) const { return std::atof(opts_.at(n).value.c_str()); }
    bool given(const std::string &n) const { return opts_.at(n).given; }
};
