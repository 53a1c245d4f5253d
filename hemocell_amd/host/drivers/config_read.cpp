// config_read -- prints values of a HemoCell XML configuration through hemo::Config
// (no device needed): config_read <file.xml> <a/b/c>...  One line per path: the
// element's text, or "<missing>" when an element on the path does not exist.
// config_read <file.xml> --xml prints the <hemocell> element as a checkpoint
// file stores it.
#include <cstdio>
#include <string>

#include "hemocell.h"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <file.xml> <a/b/c>...\n", argv[0]);
    return 2;
  }
  try {
    hemo::Config cfg(argv[1]);
    if (std::string(argv[2]) == "--xml") {
      const hemo::XMLNode* n = cfg.hemocellNode();
      if (!n) throw std::invalid_argument("no <hemocell> element");
      std::printf("%s", hemo::toXml(*n).c_str());
      return 0;
    }
    for (int a = 2; a < argc; ++a) {
      std::string path = argv[a];
      size_t p = path.find('/');
      try {
        hemo::XMLElement e = cfg[path.substr(0, p)];
        while (p != std::string::npos) {
          const size_t q = path.find('/', p + 1);
          e = e[path.substr(p + 1, q == std::string::npos ? std::string::npos : q - p - 1)];
          p = q;
        }
        std::printf("%s\n", e.getText().c_str());
      } catch (const std::invalid_argument&) {
        std::printf("<missing>\n");
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "config_read: %s\n", e.what());
    return 1;
  }
  return 0;
}
