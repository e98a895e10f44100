// facade_check.cpp -- drives the C++ facade (include/greenpeas_amd.hpp) the
// way the reference's C++ callers drive their Driver (Cli.hpp:293-302):
//
//   facade_check <max circuit> <level> [circuit ...]
//
// sizes a driver from the max circuit, compiles it (no argument), then each
// further circuit; prints each DEM as stim text between "=== <name>" lines.
// A thrown exception is printed as "EXCEPTION <class>: <message>" (exit 3).
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "greenpeas_amd.hpp"

static std::string readFile(const char *path) {
  std::ifstream f(path);
  if (!f) throw std::invalid_argument(std::string("cannot read ") + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: facade_check <max circuit> <level> [circuit ...]\n";
    return 2;
  }
  try {
    const auto level = static_cast<greenpeas_amd::CorrelationLevel>(std::stoi(argv[2]));
    auto driver = greenpeas_amd::Driver::fromStimCircuit(readFile(argv[1]), level);
    auto dem = driver.compileDetectorErrorModel();
    std::cout << "=== " << argv[1] << " errors " << dem.numErrors() << "\n" << dem.str();
    for (int i = 3; i < argc; ++i) {
      dem = driver.compileDetectorErrorModel(readFile(argv[i]));
      std::cout << "=== " << argv[i] << " errors " << dem.numErrors() << "\n" << dem.str();
    }
  } catch (const std::invalid_argument &e) {
    std::cout << "EXCEPTION invalid_argument: " << e.what() << "\n";
    return 3;
  } catch (const std::runtime_error &e) {
    std::cout << "EXCEPTION runtime_error: " << e.what() << "\n";
    return 3;
  }
  return 0;
}
