// Host utilities shared by the command-line tools.
#include <iomanip>
#include <sstream>

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


}  // namespace nnpx
