// rtps::RTPSSettings -- string -> string settings map with typed access.
//
// Same interface and semantics as the reference's RTPSSettings
// (rtpslib/RTPSSettings.h:43-107): values are stored as strings written with
// an ostringstream at its default precision (6 significant digits) and read
// back with an istringstream, so a derived float set here and read again is
// rounded to 6 digits exactly as in the reference; `changed` tracks writes.
// Header-only: used by the C++ facade and inside the library.
#pragma once

#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../rtps_c.h"

namespace rtps {

// float4 streaming as four whitespace-separated numbers (structs.h:549-553).
inline std::istream& operator>>(std::istream& is, rtps_float4& v) {
    is >> v.x >> v.y >> v.z >> v.w;
    return is;
}
inline std::ostream& operator<<(std::ostream& os, const rtps_float4& v) {
    os << v.x << " " << v.y << " " << v.z << " " << v.w;
    return os;
}

// next power of two (RTPSSettings.cpp:31-40)
inline unsigned int nlpo2(unsigned int x) {
    x--;
    x |= (x >> 1);
    x |= (x >> 2);
    x |= (x >> 4);
    x |= (x >> 8);
    x |= (x >> 16);
    return x + 1;
}

class RTPSSettings {
public:
    RTPSSettings() : changed(false) {}

    bool hasChanged() const { return changed; }
    void updated() { changed = false; }

    template <typename RT>
    RT GetSettingAs(const std::string& key, const std::string& defaultval = "0") const {
        auto it = settings.find(key);
        return typecast<RT>(it == settings.end() ? defaultval : it->second);
    }

    template <typename RT>
    void SetSetting(const std::string& key, RT value) {
        std::ostringstream oss;
        oss << value;
        settings[key] = oss.str();
        changed = true;
    }

    bool Exists(const std::string& key) const { return settings.find(key) != settings.end(); }

    void printSettings(std::ostream& os = std::cout) const {
        os << "RTPS Settings" << std::endl;
        for (const auto& kv : settings) os << kv.first << ": " << kv.second << std::endl;
    }

    const std::map<std::string, std::string>& all() const { return settings; }

private:
    std::map<std::string, std::string> settings;
    bool changed;

    template <typename RT>
    static RT typecast(const std::string& s) {
        std::istringstream is(s);
        RT v{};
        is >> v;
        return v;
    }
};

// std::string values are returned whole (the reference's istringstream >>
// string would stop at the first blank; keys read as strings here are single words).
template <>
inline std::string RTPSSettings::typecast<std::string>(const std::string& s) {
    return s;
}

// Vectors of scalars or float4 (structs.h:593-604): read until the stream ends.
template <>
inline std::vector<float> RTPSSettings::typecast<std::vector<float>>(const std::string& s) {
    std::istringstream is(s);
    std::vector<float> out;
    float v;
    while (is >> v) out.push_back(v);
    return out;
}
template <>
inline std::vector<rtps_float4> RTPSSettings::typecast<std::vector<rtps_float4>>(const std::string& s) {
    std::istringstream is(s);
    std::vector<rtps_float4> out;
    rtps_float4 v;
    while (is >> v.x >> v.y >> v.z >> v.w) out.push_back(v);
    return out;
}

}  // namespace rtps
