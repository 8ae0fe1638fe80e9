// config.txt driver over the C++ drop-in classes (include/mts.hpp): the
// reference's main (main.cc:26-133) with its command line (-i input -o
// output, appended to), configuration keys (configuration/read_config.cc),
// configuration block (configuration/write_config.h:94-127), time loops and
// error report with the interface diagnostics (helper.h:199-427,
// intersection.cc:4594-4771, douglas_adi.cc:1723-1770).
//
// Values this build evaluates on the GPU:
//   surface   = ellipsoid | torus                 the reference's default
//                                                 arguments (helper.h:144-155)
//             | ellipsoid(a,b,c,r) | torus(R,r) | bumpy(R0,A,m,k)
//   equation  = exp-sin-cos   u- = e^-t sin x sin y sin z, u+ = e^-t cos x cos y cos z
//   beta      = const(beta-,beta+) | sincos(c,beta+)   beta- = c (1 + sin x cos y / 2)
// The reference's numbered equations (0-7) and diffusion coefficients (0-4)
// and its other surfaces are not part of this build and are refused.
//
// --dry-run parses the file, writes the configuration block and builds the
// mesh and the MIB table on the host only (no GPU), then prints the crossing
// counts.
#include <getopt.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mts.hpp"

using namespace mts;

namespace {

[[noreturn]] void fatal(const std::string &msg) {
  std::cerr << "[ERROR]: " << msg << std::endl;
  std::exit(EXIT_FAILURE);
}

struct Config {
  double dom[6] = {-1, 1, -1, 1, -1, 1};  // x-min, x-max, y-min, y-max, z-min, z-max
  int n[3] = {0, 0, 0};
  double t0 = 0.0, t1 = 0.0, dt = 0.0;
  std::string surface = "ellipsoid", surface_name = "Ellipsoid";
  std::vector<double> surface_args = {2.0, 1.0, 1.0, 1.0};
  std::string method = "adi";
  std::string equation = "exp-sin-cos";
  int beta_kind = MTS_BETA_CONST;
  double beta[2] = {10.0, 1.0};
  std::string beta_text = "const(10,1)";
  int mib = 1;
  int accuracy = 2;
};

// "name(a,b,...)" -> name, {a, b, ...}
bool split_call(const std::string &v, std::string &name, std::vector<double> &args) {
  const size_t p = v.find('(');
  args.clear();
  if (p == std::string::npos) {
    name = v;
    return true;
  }
  if (v.back() != ')') return false;
  name = v.substr(0, p);
  std::stringstream ss(v.substr(p + 1, v.size() - p - 2));
  std::string item;
  while (std::getline(ss, item, ',')) {
    try {
      args.push_back(std::stod(item));
    } catch (...) {
      return false;
    }
  }
  return true;
}

// configuration/read_config.cc:16-100: "key = value" lines, '#' lines and
// empty lines skipped, spaces removed.
Config read_config(const std::string &file) {
  std::ifstream in(file);
  if (!in.good()) fatal("Read configuration file failed, check path of " + file);
  Config c;
  std::string line;
  while (std::getline(in, line)) {
    if (line.find('#') != std::string::npos || line.empty()) continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key = line.substr(0, eq), val = line.substr(eq + 1);
    key.erase(std::remove(key.begin(), key.end(), ' '), key.end());
    val.erase(std::remove(val.begin(), val.end(), ' '), val.end());
    val.erase(std::remove(val.begin(), val.end(), '\r'), val.end());
    try {
      // the domain bounds are read as integers, as the reference reads them
      // (stoi into doubles, configuration/read_config.cc:39-56): "3.99" is 3
      if (key == "x-min") c.dom[0] = std::stoi(val);
      else if (key == "x-max") c.dom[1] = std::stoi(val);
      else if (key == "y-min") c.dom[2] = std::stoi(val);
      else if (key == "y-max") c.dom[3] = std::stoi(val);
      else if (key == "z-min") c.dom[4] = std::stoi(val);
      else if (key == "z-max") c.dom[5] = std::stoi(val);
      else if (key == "x-grid") c.n[0] = std::stoi(val);
      else if (key == "y-grid") c.n[1] = std::stoi(val);
      else if (key == "z-grid") c.n[2] = std::stoi(val);
      else if (key == "time-start") c.t0 = std::stod(val);
      else if (key == "time-finish") c.t1 = std::stod(val);
      else if (key == "time-step") c.dt = std::stod(val);
      else if (key == "surface") {
        std::string name;
        std::vector<double> args;
        if (!split_call(val, name, args)) fatal("cannot read surface \"" + val + "\"");
        if (name == "ellipsoid") {
          c.surface_name = "Ellipsoid";
          c.surface_args = args.empty() ? std::vector<double>{2.0, 1.0, 1.0, 1.0} : args;
          if (c.surface_args.size() != 4) fatal("ellipsoid takes (a,b,c,r)");
        } else if (name == "torus") {
          c.surface_name = "Torus";
          c.surface_args = args.empty() ? std::vector<double>{1.5, 0.7} : args;
          if (c.surface_args.size() != 2) fatal("torus takes (R,r)");
        } else if (name == "bumpy") {
          c.surface_name = "Bumpy-Sphere";
          c.surface_args = args;
          if (c.surface_args.size() != 4) fatal("bumpy takes (R0,A,m,k)");
        } else if (name == "tanglecube" || name == "cube" || name == "cylinder" || name == "cone" ||
                   name == "pile" || name == "dupin_cyclide" || name == "molecular" ||
                   name == "heart") {
          fatal("surface \"" + name + "\" is not available in this build: use ellipsoid, torus "
                "or bumpy, or a Surface_Cartesian subclass through the C++ classes");
        } else {
          fatal("Surface type not found! check configuration file");
        }
        c.surface = name;
      } else if (key == "temporal-method") {
        if (val != "adi" && val != "lod-ie" && val != "lod-cn" && val != "ts")
          fatal("Temporal-method type not found! Currently support: \"adi\", \"lod-ie\", "
                "\"lod-cn\", \"ts\", check configuration file");
        c.method = val;
      } else if (key == "equation") {
        if (val != "exp-sin-cos")
          fatal("equation \"" + val + "\" is not available in this build: the kernels evaluate "
                "exp-sin-cos (u- = e^-t sin x sin y sin z, u+ = e^-t cos x cos y cos z)");
        c.equation = val;
      } else if (key == "beta") {
        std::string name;
        std::vector<double> args;
        if (!split_call(val, name, args) || args.size() != 2 ||
            (name != "const" && name != "sincos"))
          fatal("beta \"" + val + "\" is not available in this build: use const(beta-,beta+) "
                "or sincos(c,beta+)");
        c.beta_kind = name == "const" ? MTS_BETA_CONST : MTS_BETA_SINCOS;
        c.beta[0] = args[0];
        c.beta[1] = args[1];
        c.beta_text = val;
      } else if (key == "spatial-method") {
        if (val == "mib-v1") c.mib = 1;
        else if (val == "mib-v2") c.mib = 2;
        else fatal("Spatial method type not found! Currently support: \"mib-v1\", \"mib-v2\", "
                   "check configuration file");
      } else if (key == "spatial-accuracy") {
        c.accuracy = std::stoi(val);
        if (c.accuracy != 2) fatal("only 2nd-order spatial accuracy is available in this build");
      }
    } catch (const std::logic_error &) {
      fatal("cannot read the value of " + key);
    }
  }
  for (int a = 0; a < 3; a++)
    if (c.n[a] < 5) fatal("x-grid, y-grid and z-grid need at least 5 nodes");
  if (!(c.dt > 0.0)) fatal("time-step must be positive");
  return c;
}

const char *method_name(const std::string &m) {
  return m == "adi" ? "ADI" : m == "lod-ie" ? "LOD-IE" : m == "lod-cn" ? "LOD-CN" : "TS";
}

// configuration/write_config.h:94-127
void write_configuration(std::ofstream &out, const Config &c) {
  out << "----------------- Configuration informations -----------------" << std::endl;
  out << "Mesh Size                          :   "
      << "NX = " << c.n[0] << "; NY = " << c.n[1] << "; NZ = " << c.n[2] << std::endl;
  out << "Time Step                          :   " << std::setprecision(1) << std::scientific
      << c.dt << std::endl;
  out << "Jump                               :   " << std::fixed << "approximate" << std::endl;
  out << "Diffusion Coefficient Type         :   " << c.beta_text << std::endl;
  out << "Equation Type                      :   " << c.equation << std::endl;
  out << "Order of Spatial Accuracy          :   " << c.accuracy << std::endl;
  out << "Surface                            :   " << c.surface_name << std::endl;
  out << "Temporal Method                    :   " << method_name(c.method) << std::endl;
  out << "Spatial Method                     :   " << (c.mib == 1 ? "MIB-V1" : "MIB-V2")
      << std::endl
      << std::endl;
}

std::unique_ptr<Surface_Cartesian> make_surface(const Config &c) {
  if (c.surface == "ellipsoid") return std::unique_ptr<Surface_Cartesian>(new Surface_Ellipsoid(c.surface_args));
  if (c.surface == "torus") return std::unique_ptr<Surface_Cartesian>(new Surface_Torus(c.surface_args));
  return std::unique_ptr<Surface_Cartesian>(new Surface_Bumpy(c.surface_args));
}

// Write_txt (helper.h:442-575), which the reference's main calls after every
// run: for each crossing (z lines, then y, then x, each in the reference's
// loop order) the node its ID's parity picks (left_loc for odd |ID|, else
// left_loc + 1), with its value and its error against the outer exact
// solution at the configuration's finish time, into "Numerical
// Solutions.txt" and "Numerical Error.txt" in the working directory.  The
// reference's Equation pointer there dangles (it outlives the object it
// points to); here the equation is the run's.
void write_txt(const Intersections &inter, const Mesh &mesh, Beta &beta,
               const CubicDoub &uh, double t) {
  std::ofstream sol("Numerical Solutions.txt", std::ios::out);
  std::ofstream err("Numerical Error.txt", std::ios::out);
  if (!sol.good() || !err.good()) fatal("No such file is found!");
  sol << std::setiosflags(std::ios::right) << std::setprecision(6) << std::fixed;
  err << std::setiosflags(std::ios::right) << std::setprecision(6) << std::fixed;
  const Equation eq(t, beta);
  auto one = [&](const mts_crossing &c, int ix, int iy, int iz) {
    const int node = (std::abs(c.id) % 2 == 1) ? c.left_loc : c.left_loc + 1;
    if (c.dir == 2) iz = node;
    else if (c.dir == 1) iy = node;
    else ix = node;
    const double v = uh[ix][iy][iz];
    const double e = std::abs(v - eq.Outer_u(mesh.xi[ix], mesh.yi[iy], mesh.zi[iz]));
    sol << std::setw(5) << ix << std::setw(5) << iy << std::setw(5) << iz << std::setw(20) << v
        << std::endl;
    err << std::setw(5) << ix << std::setw(5) << iy << std::setw(5) << iz << std::setw(20) << e
        << std::endl;
  };
  for (int ix = 1; ix < mesh.nx - 1; ix++)
    for (int iy = 1; iy < mesh.ny - 1; iy++)
      for (const mts_crossing &c : inter.ifpz[ix][iy]) one(c, ix, iy, 0);
  for (int ix = 1; ix < mesh.nx - 1; ix++)
    for (int iz = 1; iz < mesh.nz - 1; iz++)
      for (const mts_crossing &c : inter.ifpy[ix][iz]) one(c, ix, 0, iz);
  for (int iy = 1; iy < mesh.ny - 1; iy++)
    for (int iz = 1; iz < mesh.nz - 1; iz++)
      for (const mts_crossing &c : inter.ifpx[iy][iz]) one(c, 0, iy, iz);
}

void print_help(const char *bin) {
  std::cout << "Usage: " << bin << " -i input -o output [--dry-run]\n\n"
            << "-i, --input [required]\n      input configuration file\n"
            << "-o, --output [required]\n      output file (appended to)\n"
            << "--dry-run\n      parse and set up on the host only\n";
}

}  // namespace

int main(int argc, char *argv[]) {
  std::cout << "[INFO]: PROGRAM STARTS ... " << std::endl;
  const auto t_start = std::chrono::high_resolution_clock::now();
  std::string input, output;
  bool dry = false;
  static struct option opts[] = {{"input", required_argument, nullptr, 'i'},
                                 {"output", required_argument, nullptr, 'o'},
                                 {"help", no_argument, nullptr, 'h'},
                                 {"dry-run", no_argument, nullptr, 'd'},
                                 {nullptr, no_argument, nullptr, 0}};
  int o;
  while ((o = getopt_long(argc, argv, "i:o:h", opts, nullptr)) != -1) {
    switch (o) {
      case 'i':
        input = optarg;
        std::cout << "[INFO]: reading configuration from file: " << input << std::endl;
        break;
      case 'o':
        output = optarg;
        std::cout << "[INFO]: writing result to file: " << output << std::endl;
        break;
      case 'd':
        dry = true;
        break;
      case 'h':
        print_help(argv[0]);
        return EXIT_SUCCESS;
      default:
        print_help(argv[0]);
        return EXIT_FAILURE;
    }
  }
  if (input.empty() || output.empty()) {
    print_help(argv[0]);
    return EXIT_FAILURE;
  }
  set_exit_on_error(true);  // the reference's command-line behaviour
  const Config c = read_config(input);
  std::ofstream out(output, std::ios::out | std::ios::app);
  if (!out.good()) fatal("I/O failure when opening " + output);
  write_configuration(out, c);

  std::unique_ptr<Surface_Cartesian> surf = make_surface(c);
  Beta beta(c.beta_kind, c.beta[0], c.beta[1]);
  const VecDoub domain(c.dom, c.dom + 6);
  const VecInt size(c.n, c.n + 3);
  Mesh mesh(domain, size, *surf);
  Intersections inter(*surf, mesh, beta, c.accuracy, c.mib, out);
  if (dry) {
    size_t nc[3] = {0, 0, 0};
    const Intersections::Cubic_Intersection_Data *ifp[3] = {&inter.ifpx, &inter.ifpy, &inter.ifpz};
    for (int d = 0; d < 3; d++)
      for (auto &row : *ifp[d])
        for (auto &line : row) nc[d] += line.size();
    std::cout << "[INFO]: crossings x " << nc[0] << " y " << nc[1] << " z " << nc[2] << std::endl;
    std::cout << "[INFO]: PROGRAM FINISHED!" << std::endl;
    return EXIT_SUCCESS;
  }

  const VecDoub time = {c.t0, c.t1, c.dt};
  CubicDoub uh;
  // The time loops of helper.h:199-427: loop = (int)(t_finish / dt), t += dt.
  const int loop = (int)(c.t1 / c.dt);
  double t_now = c.t0;
  // the report of helper.h:222-236: time, fictitious-point and jump errors,
  // then the solution's error
  auto report = [&](detail::Solver &s, const char *title) {
    out << title << std::endl << std::endl;
    out << std::setprecision(1) << std::scientific << "T = " << t_now << std::endl;
    out << std::fixed;
    Equation eq(t_now, beta);
    inter.Refresh_Fp(eq);
    inter.Error_Fp(out);
    inter.Refresh_Jump(eq, uh, beta);
    inter.Error_Jump(out);
    s.Error(eq, uh, out);
  };
  if (c.method == "adi") {
    Equation eq0(c.t0, beta);
    Douglas_ADI adi(inter, mesh, beta, time, c.accuracy);
    adi.set_host_sync(false);
    adi.Initialization(eq0, uh);
    for (int j = 0; j < loop; j++) {
      t_now += c.dt;
      Equation eq(t_now, beta);
      adi.Solve_2nd(eq, inter, uh, beta);
    }
    report(adi, "------------------ Error of Douglas-ADI with MIB ------------------");
  } else if (c.method == "lod-ie" || c.method == "lod-cn") {
    char m = c.method == "lod-ie" ? 'I' : 'C';
    Equation eq0(c.t0, beta);
    LOD lod(inter, mesh, beta, time);
    lod.set_host_sync(false);
    lod.Initialization(eq0, uh);
    for (int j = 0; j < loop; j++) {
      t_now += c.dt;
      Equation eq(t_now, beta);
      lod.Solve_2nd(eq, inter, uh, beta, m);
    }
    report(lod, "------------------- Error of LOD with MIB --------------------");
  } else {
    Equation eq0(c.t0, beta);
    TS ts(inter, mesh, beta, time);
    ts.set_host_sync(false);
    ts.Initialization(eq0, uh);
    for (int j = 0; j < loop; j++) {
      const double tn = t_now;
      t_now += c.dt;
      Equation e_n(tn, beta), e_half(tn + 0.5 * c.dt, beta), e_np1(t_now, beta);
      ts.Solve_2nd(e_n, e_half, e_np1, inter, uh, beta);
    }
    report(ts, "---------------------- Error of TS with MIB ----------------------");
  }
  const std::chrono::duration<double> period = std::chrono::high_resolution_clock::now() - t_start;
  out << std::setprecision(1) << std::endl;
  out << "CPU time cost: " << period.count() << " seconds" << std::endl << std::endl;
  write_txt(inter, mesh, beta, uh, c.t1);
  std::cout << "[INFO]: PROGRAM FINISHED!" << std::endl;
  return EXIT_SUCCESS;
}
