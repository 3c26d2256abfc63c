// main.cpp -- command-line driver with the reference's behaviour (src/main.cpp:9-39): read
// ./config/settings.cfg (or the path given as the first argument), build the equation and
// the solver, run the case's time loop.  One process drives one GPU; there is no MPI.
#include <iostream>
#include <map>
#include <string>

#include "Equations.hpp"
#include "ReadConfig.hpp"
#include "WENOFV.hpp"

int main(int argc, char *argv[]) {
    const std::string inputfile = argc > 1 ? argv[1] : "./config/settings.cfg";
    try {
        TwoPhaseEquation equation;
        std::cout << "Starting to read configuration file......" << std::endl;
        CConfig config(inputfile);
        std::map<std::string, std::string> option;
        config.read(option);
        CWENOFV solver(&equation, option);
        solver.run();
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
