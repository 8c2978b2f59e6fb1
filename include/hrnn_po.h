// hrnn_po.h -- the part of Boost.Program_options the HRNN drivers use, for
// builds without Boost.  Namespace po, as the reference's include/utils.h:20-21
// names it (`namespace po = boost::program_options;`), so the reference's
// src/hrnn_build.cpp, src/hrnn_search.cpp and src/hrnn_dynamic_test.cpp
// compile unchanged against this repository's include/ directory.
//
// Covered (the interface the three drivers use, from Boost's documented
// behaviour): options_description with add_options()(name[,short], [value],
// help); value<T>(&var) with ->default_value(v) and ->required() for
// T = std::string, int, double (and other stream-readable types);
// parse_command_line (--name value, --name=value, -s value, -svalue);
// variables_map with count(); store (values checked and defaulted);
// notify (required options checked, bound variables assigned); printing an
// options_description.  Errors are po::error (a std::logic_error), as Boost's.
#ifndef HRNN_PO_H
#define HRNN_PO_H

#include <algorithm>
#include <cstddef>
#include <map>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace po {

class error : public std::logic_error {
 public:
    explicit error(const std::string &what) : std::logic_error(what) {}
};

// What an option does with its token: check it, keep it, assign it to the
// caller's variable; a default; whether it must be given.
class value_semantic {
 public:
    virtual ~value_semantic() = default;
    virtual void check(const std::string &option, const std::string &token) const = 0;
    virtual void assign(const std::string &token) const = 0;
    virtual bool has_default() const = 0;
    virtual void assign_default() const = 0;
    virtual std::string default_text() const = 0;
    bool is_required() const { return required_; }

 protected:
    bool required_ = false;
};

namespace detail {
template <class T>
bool read_token(const std::string &s, T &out) {
    std::istringstream in(s);
    in >> out;
    if (in.fail()) return false;
    in >> std::ws;
    return in.eof();
}
inline bool read_token(const std::string &s, std::string &out) {
    out = s;
    return true;
}
template <class T>
std::string show(const T &v) {
    std::ostringstream o;
    o << v;
    return o.str();
}
}  // namespace detail

template <class T>
class typed_value : public value_semantic {
 public:
    explicit typed_value(T *store) : store_(store) {}
    typed_value *default_value(const T &v) {
        def_.reset(new T(v));
        return this;
    }
    typed_value *required() {
        required_ = true;
        return this;
    }
    void check(const std::string &option, const std::string &token) const override {
        T tmp;
        if (!detail::read_token(token, tmp))
            throw error("the argument ('" + token + "') for option '--" + option + "' is invalid");
    }
    void assign(const std::string &token) const override {
        T tmp;
        detail::read_token(token, tmp);
        if (store_) *store_ = tmp;
    }
    bool has_default() const override { return def_ != nullptr; }
    void assign_default() const override {
        if (store_ && def_) *store_ = *def_;
    }
    std::string default_text() const override { return def_ ? detail::show(*def_) : std::string(); }

 private:
    T *store_;
    std::unique_ptr<T> def_;
};

template <class T>
typed_value<T> *value(T *store) {
    return new typed_value<T>(store);
}
template <class T>
typed_value<T> *value() {
    return new typed_value<T>(nullptr);
}

struct option_description {
    std::string long_name, short_name, help;
    std::shared_ptr<value_semantic> semantic;   // null: a switch without a value
};

class options_description {
 public:
    explicit options_description(std::string caption = std::string()) : caption_(std::move(caption)) {}

    class easy_init {
     public:
        explicit easy_init(options_description *o) : o_(o) {}
        easy_init &operator()(const char *name, const char *help) {
            o_->add(name, nullptr, help);
            return *this;
        }
        easy_init &operator()(const char *name, value_semantic *s, const char *help) {
            o_->add(name, s, help);
            return *this;
        }
        easy_init &operator()(const char *name, value_semantic *s) {
            o_->add(name, s, "");
            return *this;
        }

     private:
        options_description *o_;
    };
    easy_init add_options() { return easy_init(this); }

    // The option named `token` (long name, or short name when is_short).
    const option_description *find(const std::string &token, bool is_short) const {
        for (auto &o : opts_)
            if ((is_short ? o->short_name : o->long_name) == token) return o.get();
        return nullptr;
    }
    const std::vector<std::shared_ptr<option_description>> &options() const { return opts_; }

    friend std::ostream &operator<<(std::ostream &os, const options_description &d) {
        if (!d.caption_.empty()) os << d.caption_ << ":\n";
        std::vector<std::string> left;
        size_t width = 0;
        for (auto &o : d.opts_) {
            std::string s = "  ";
            s += o->short_name.empty() ? "--" + o->long_name : "-" + o->short_name + " [ --" + o->long_name + " ]";
            if (o->semantic) {
                s += " arg";
                if (o->semantic->has_default()) s += " (=" + o->semantic->default_text() + ")";
            }
            width = std::max(width, s.size());
            left.push_back(s);
        }
        for (size_t i = 0; i < d.opts_.size(); i++)
            os << left[i] << std::string(width + 2 - left[i].size(), ' ') << d.opts_[i]->help << "\n";
        return os;
    }

 private:
    void add(const char *name, value_semantic *s, const char *help) {
        auto o = std::make_shared<option_description>();
        std::string n(name);
        const size_t comma = n.find(',');
        o->long_name = n.substr(0, comma);
        if (comma != std::string::npos) o->short_name = n.substr(comma + 1);
        o->help = help ? help : "";
        o->semantic.reset(s);
        opts_.push_back(o);
    }
    std::string caption_;
    std::vector<std::shared_ptr<option_description>> opts_;
};

// One recognised option of the command line and its token.
struct parsed_option {
    const option_description *desc;
    std::string token;
};
struct parsed_options {
    const options_description *description = nullptr;
    std::vector<parsed_option> options;
};

inline parsed_options parse_command_line(int argc, const char *const argv[],
                                         const options_description &desc) {
    parsed_options out;
    out.description = &desc;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        std::string name, token;
        bool is_short = false, inline_token = false;
        if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                token = name.substr(eq + 1);
                name = name.substr(0, eq);
                inline_token = true;
            }
        } else if (a.size() >= 2 && a[0] == '-' && a[1] != '-') {
            is_short = true;
            name = a.substr(1, 1);
            if (a.size() > 2) {
                token = a.substr(2);
                inline_token = true;
            }
        } else {
            throw error("too many positional options have been specified on the command line");
        }
        const option_description *o = desc.find(name, is_short);
        if (!o) throw error("unrecognised option '" + a + "'");
        if (o->semantic) {
            if (!inline_token) {
                if (i + 1 >= argc) throw error("the required argument for option '--" + o->long_name + "' is missing");
                token = argv[++i];
            }
        } else if (inline_token) {
            throw error("option '--" + o->long_name + "' does not take any arguments");
        }
        out.options.push_back({o, token});
    }
    return out;
}

class variable_value {
 public:
    variable_value() = default;
    variable_value(std::string token, bool defaulted) : token_(std::move(token)), defaulted_(defaulted) {}
    bool defaulted() const { return defaulted_; }
    const std::string &token() const { return token_; }

 private:
    std::string token_;
    bool defaulted_ = false;
};

class variables_map : public std::map<std::string, variable_value> {
 public:
    const options_description *description = nullptr;
};

inline void store(const parsed_options &p, variables_map &vm) {
    vm.description = p.description;
    for (auto &po_ : p.options) {
        if (po_.desc->semantic) po_.desc->semantic->check(po_.desc->long_name, po_.token);
        vm[po_.desc->long_name] = variable_value(po_.token, false);   // the last occurrence wins
    }
    if (!p.description) return;
    for (auto &o : p.description->options())
        if (o->semantic && o->semantic->has_default() && !vm.count(o->long_name))
            vm[o->long_name] = variable_value(o->semantic->default_text(), true);
}

inline void notify(variables_map &vm) {
    if (!vm.description) return;
    for (auto &o : vm.description->options()) {
        if (!o->semantic) continue;
        auto it = vm.find(o->long_name);
        if (it == vm.end() || it->second.defaulted()) {
            if (o->semantic->is_required())
                throw error("the option '--" + o->long_name + "' is required but missing");
            o->semantic->assign_default();
        } else {
            o->semantic->assign(it->second.token());
        }
    }
}

}  // namespace po

#endif  // HRNN_PO_H
