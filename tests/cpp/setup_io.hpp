// setup_io.hpp -- reader of the named-array files written by tests/setup_io.py (test code).
#ifndef S2L_TEST_SETUP_IO_HPP
#define S2L_TEST_SETUP_IO_HPP

#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace s2ltest {

struct Record {
    int type = 0;
    std::vector<unsigned char> bytes;
    size_t count = 0;
};

class ArrayFile {
public:
    explicit ArrayFile(const std::string& path)
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open " + path);
        char magic[8];
        f.read(magic, 8);
        if (std::memcmp(magic, "S2LSETUP", 8) != 0) throw std::runtime_error("not a setup file: " + path);
        uint32_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 4);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t len = 0;
            f.read(reinterpret_cast<char*>(&len), 4);
            std::string name(len, '\0');
            f.read(&name[0], len);
            uint8_t type = 0;
            uint64_t count = 0;
            f.read(reinterpret_cast<char*>(&type), 1);
            f.read(reinterpret_cast<char*>(&count), 8);
            static const size_t esz[4] = {4, 8, 8, 1};
            if (type > 3) throw std::runtime_error("bad record type in " + path);
            Record r;
            r.type = type;
            r.count = count;
            r.bytes.resize(count * esz[type]);
            f.read(reinterpret_cast<char*>(r.bytes.data()), static_cast<std::streamsize>(r.bytes.size()));
            if (!f) throw std::runtime_error("truncated setup file " + path);
            rec_[name] = std::move(r);
        }
    }
    bool has(const std::string& k) const { return rec_.count(k) != 0; }
    template <typename T> std::vector<T> get(const std::string& k, int type) const
    {
        auto it = rec_.find(k);
        if (it == rec_.end()) throw std::runtime_error("missing record " + k);
        if (it->second.type != type) throw std::runtime_error("record " + k + " has another type");
        std::vector<T> v(it->second.count);
        std::memcpy(v.data(), it->second.bytes.data(), it->second.bytes.size());
        return v;
    }
    std::vector<int32_t> i32(const std::string& k) const { return get<int32_t>(k, 0); }
    std::vector<int64_t> i64(const std::string& k) const { return get<int64_t>(k, 1); }
    std::vector<double> f64(const std::string& k) const { return get<double>(k, 2); }
    int64_t scalar(const std::string& k) const { return i64(k).at(0); }
    std::string text(const std::string& k) const
    {
        auto it = rec_.find(k);
        if (it == rec_.end() || it->second.type != 3) throw std::runtime_error("missing text " + k);
        return std::string(it->second.bytes.begin(), it->second.bytes.end());
    }

private:
    std::map<std::string, Record> rec_;
};

} // namespace s2ltest

#endif
