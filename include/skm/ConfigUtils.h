// skm/ConfigUtils.h -- the SKMReg parameter set with the reference's keys and
// defaults (include/ConfigUtils.h:29-62) and its "key value" command-line
// overrides (src/SKMReg.cxx:591-596), without a JSON dependency.
//
// Numeric options are per-scale arrays; CURRENT(option, scale) clamps the
// scale to the last entry like the reference's CURRENT macro (SKMReg.cxx:48-49).
#pragma once

#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace skm {

class Config {
public:
    Config()
    {
        text_ = {{"reference_filename", "/tmp/ref.vtk"}, {"target_filename", "/tmp/tar.vtk"},
                 {"reference_landmarks", "none"},        {"target_landmarks", "none"},
                 {"weight_filename", "none"},            {"weight_tensor_filename", "none"},
                 {"temp_directory", "/tmp/skm/"},        {"metric", "mse"}};
        num_ = {{"sigma_C4", {64, 32, 16}},        {"sigma_C0", {128, 64, 32}},
                {"cp_spacing_scaler", {4.0}},      {"tensor_weights", {400, 200, 100}},
                {"reg_l1", {0, 0, 0}},             {"reg_l21", {0, 0, 0}},
                {"reg_l2", {0, 0, 0}},             {"reg_rkhs", {0, 0, 0}},
                {"reg_rd", {0, 0, 0}},             {"reg_rd_scaling", {1, 1, 1}},
                {"reg_pg", {0, 0, 0}},             {"reg_pg_scaling", {1, 1, 1}},
                {"num_scales", {3}},               {"restart_at", {0}},
                {"num_function_evaluations", {500, 500, 500}},
                {"initial_step_size", {0.1, 0.1, 0.1}},
                {"sampling_rate", {5, 4, 3}},      {"neighborhood_sampling", {15, 12, 9}},
                {"intensity_thresholds", {-1500, 3500}},
                {"smoothing_variance", {2, 1, 0}}, {"verbosity", {1}}};
    }

    double CURRENT(const std::string &option, unsigned scale) const
    {
        const auto &v = num_.at(option);
        return v[scale < v.size() ? scale : v.size() - 1];
    }
    const std::vector<double> &numbers(const std::string &option) const { return num_.at(option); }
    const std::string &text(const std::string &option) const { return text_.at(option); }

    // Applies "key value" pairs; a value is a number, a bracketed list
    // "[a, b, c]" or a string (quotes are stripped).
    void apply_argv(int argc, const char *const *argv)
    {
        for (int i = 1; i + 1 < argc; i += 2) set(argv[i], argv[i + 1]);
    }

    void set(const std::string &key, std::string value)
    {
        std::string v;
        for (char c : value)
            if (c != '"') v += c;
        if (num_.count(key)) {
            std::vector<double> out;
            std::string body = v;
            if (!body.empty() && body.front() == '[') body = body.substr(1, body.size() - 2);
            std::stringstream ss(body);
            std::string tok;
            while (std::getline(ss, tok, ',')) out.push_back(std::stod(tok));
            if (out.empty()) throw std::invalid_argument("empty value for " + key);
            num_[key] = out;
        } else if (text_.count(key)) {
            text_[key] = v;
        } else {
            throw std::invalid_argument("unknown option " + key);
        }
    }

private:
    std::map<std::string, std::string> text_;
    std::map<std::string, std::vector<double>> num_;
};

}  // namespace skm
