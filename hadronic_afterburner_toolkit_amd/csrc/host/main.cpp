// hbt_afterburner.e -- command-line driver of the HBT path, the
// analyze_HBT=1 branch of the reference's hadronic_afterburner_tools.e
// (src/main.cpp:21-46, Analysis::HBTAnalysis Analysis.cpp:817-835):
// reads ./parameters.dat and key=value overrides, samples from
// results/particle_samples.gz (read_in_mode 10) or results/particle_list.bin
// (read_in_mode 9), writes the HBT files to results/.
#include <chrono>
#include <iostream>
#include <memory>
#include <string>

#include "HBT_correlation.h"
#include "ParameterReader.h"
#include "Random.h"
#include "particleSamples.h"

int main(int argc, char *argv[]) {
    std::cout << "HBT pair correlations on MI355X (hadronic_afterburner_toolkit HBT path)"
              << std::endl;
    const std::string path = "results";
    const auto t0 = std::chrono::steady_clock::now();
    ParameterReader paraRdr;
    paraRdr.readFromFile("parameters.dat");
    paraRdr.readFromArguments(argc, argv);
    paraRdr.echo();
    if (paraRdr.getVal("analyze_HBT") != 1) {
        std::cout << "analyze_HBT is not 1: nothing to do (this build runs the HBT path only)"
                  << std::endl;
        return 0;
    }
    // Analysis::InitializeAnalysis (Analysis.cpp:49-52)
    const int randomSeed = (int)paraRdr.getVal("randomSeed", -1);
    auto ran_gen_ptr = std::make_shared<RandomUtil::Random>(randomSeed);
    auto particle_list = std::make_shared<particleSamples>(paraRdr, path, ran_gen_ptr);
    HBT_correlation HBT_analysis(paraRdr, path, ran_gen_ptr);
    int event_id = 0;
    while (!particle_list->end_of_file()) {
        std::cout << "Reading event: " << event_id + 1 << " ... " << std::endl;
        particle_list->read_in_particle_samples_and_filter();
        particle_list->read_in_particle_samples_mixed_event_and_filter();
        const int nev = particle_list->get_number_of_events();
        std::cout << "nev = " << nev << std::endl;
        HBT_analysis.calculate_HBT_correlation_function(particle_list);
        event_id += nev;
    }
    HBT_analysis.output_HBTcorrelation();
    const double dt =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Program totally finished in " << dt << " sec." << std::endl;
    return 0;
}
