// ReadConfig.hpp -- reader of config/settings.cfg: `key = value` lines, '#' comments,
// blank and CR lines skipped, keys and values trimmed of spaces and a trailing CR (the
// behaviour of the reference's CConfig::read, include/ReadConfig.hpp:15-56, and
// sc_common::trim, third_party/scTools/Common.h:91-107).  A missing file is an error the
// caller sees (an exception), not a process exit.
#pragma once
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

class CConfig {
public:
    std::string m_inputFileName;
    explicit CConfig(const std::string &str) : m_inputFileName(str) {}

    static void trim(std::string &str) {
        const size_t first = str.find_first_not_of(' ');
        if (first == std::string::npos) {
            str.clear();
            return;
        }
        size_t last = str.find_last_not_of(' ');
        if (str[last] == '\r') last--;
        str = str.substr(first, last - first + 1);
    }

    void read(std::map<std::string, std::string> &option) const {
        std::ifstream input(m_inputFileName);
        if (!input.is_open()) throw std::runtime_error("Failed to open file" + m_inputFileName);
        std::string line;
        while (std::getline(input, line)) {
            if (line.empty() || line[0] == '#' || line[0] == '\r') continue;
            const size_t eq = line.find('=');
            if (eq == std::string::npos) {
                std::cerr << "Invalid line: " << line << std::endl;
                continue;
            }
            std::string key = line.substr(0, eq), value = line.substr(eq + 1);
            trim(key);
            trim(value);
            option[key] = value;
        }
    }
};
