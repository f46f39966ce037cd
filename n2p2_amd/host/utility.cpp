// Host utilities shared by the command-line tools.
#include <cstdarg>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "nnpx_mode.hpp"

namespace nnpx {

// Column header block of n2p2's output files (utility.cpp:111-204 layout).
void writeFileHeader(std::ofstream& o, std::vector<std::string> const& title, std::vector<size_t> const& size,
                  std::vector<std::string> const& name, std::vector<std::string> const& info)
{
    const size_t std_width = 80;
    const std::string sep = std::string(std_width, '#') + "\n";
    o << sep;
    for (auto const& t : title) o << "# " << t << "\n";
    o << sep;
    size_t wn = 0;
    for (auto const& n : name) wn = std::max(wn, n.size());
    std::ostringstream s;
    s << "# " << std::left << std::setw(4) << "Col" << ' ' << std::left << std::setw(wn) << "Name" << ' '
      << "Description\n";
    o << s.str() << sep;
    for (size_t i = 0; i < name.size(); ++i) {
        std::ostringstream r;
        r << "# " << std::left << std::setw(4) << i + 1 << ' ' << std::left << std::setw(wn) << name[i] << ' '
          << info[i] << '\n';
        o << r.str();
    }
    size_t wd = 0;
    for (size_t c : size) wd += c + 1;
    wd -= 1;
    o << (wd > std_width ? std::string(wd, '#') + "\n" : sep);
    std::ostringstream num, nam;
    num << '#';
    nam << '#';
    for (size_t i = 0; i < size.size(); ++i) {
        size_t n = size[i] - (i == 0 ? 2 : 0);
        std::string nm = name[i];
        if (nm.size() > n) nm.erase(n);
        num << ' ' << std::right << std::setw(n) << i + 1;
        nam << ' ' << std::right << std::setw(n) << nm;
    }
    o << num.str() << '\n' << nam.str() << '\n' << std::string(wd, '#') << "\n";
}


std::string strpr(const char* format, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, format);
    vsnprintf(buf, sizeof buf, format, ap);
    va_end(ap);
    return buf;
}

std::vector<std::vector<std::string>> readStructureBlocks(std::string const& fileName)
{
    std::ifstream in(fileName);
    if (!in) throw std::runtime_error("ERROR: Could not open file: \"" + fileName + "\".\n");
    std::vector<std::vector<std::string>> blocks;
    std::vector<std::string> cur;
    std::string line;
    bool open = false;
    while (std::getline(in, line)) {
        std::istringstream w(line);
        std::string k;
        if (!(w >> k)) continue;
        if (k == "begin") {
            cur.clear();
            open = true;
        }
        if (open) cur.push_back(line);
        if (k == "end" && open) {
            blocks.push_back(cur);
            open = false;
        }
    }
    return blocks;
}

bool settingsFileHasKeyword(std::string const& fileName, std::string const& keyword)
{
    std::ifstream in(fileName);
    std::string line;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.erase(h);
        std::istringstream w(line);
        std::string k;
        if (w >> k && k == keyword) return true;
    }
    return false;
}

void PackedStructures::add(Structure const& s, double convLength)
{
    for (auto const& a : s.atoms) {
        for (int c = 0; c < 3; ++c) position.push_back(a.r[c] * convLength);
        element.push_back((int)a.element);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) box.push_back(s.box[r][c] * convLength);
    periodic.push_back(s.isPeriodic ? 1 : 0);
    offset.push_back(offset.back() + (int64_t)s.numAtoms);
}

void PackedStructures::clear()
{
    offset.assign(1, 0);
    position.clear();
    box.clear();
    element.clear();
    periodic.clear();
}

}  // namespace nnpx
