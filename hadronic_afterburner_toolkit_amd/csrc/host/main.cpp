// hbt_afterburner.e -- command-line driver of the HBT path and the balance
// function, the analyze_HBT=1 and analyze_balance_function=1 branches of
// the reference's hadronic_afterburner_tools.e (src/main.cpp:21-46,
// Analysis::HBTAnalysis Analysis.cpp:817-835, Analysis::
// BalanceFunctionAnalysis :856-874, in that order, sharing one random
// stream): reads ./parameters.dat and key=value overrides, samples from
// results/particle_samples.gz (read_in_mode 10) or results/particle_list.bin
// (read_in_mode 9), writes the files to results/.
#include <chrono>
#include <iostream>
#include <memory>
#include <string>

#include "BalanceFunction.h"
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
    const bool hbt = paraRdr.getVal("analyze_HBT", 0) == 1;
    const bool bf = paraRdr.getVal("analyze_balance_function", 0) == 1;
    if (!hbt && !bf) {
        std::cout << "neither analyze_HBT nor analyze_balance_function is 1: nothing to do "
                     "(this executable runs those two analyses)"
                  << std::endl;
        return 0;
    }
    // Analysis::InitializeAnalysis (Analysis.cpp:49-52)
    const int randomSeed = (int)paraRdr.getVal("randomSeed", -1);
    auto ran_gen_ptr = std::make_shared<RandomUtil::Random>(randomSeed);
    if (hbt) {
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
    }
    if (bf) {
        // the balance-function lists are filtered from all species
        ParameterReader all = paraRdr;
        all.setVal("particle_monval", 99999);
        auto particle_list = std::make_shared<particleSamples>(all, path, ran_gen_ptr);
        BalanceFunction BF_analysis(paraRdr, path, ran_gen_ptr);
        int event_id = 0;
        while (!particle_list->end_of_file()) {
            std::cout << "Reading event: " << event_id + 1 << " ... " << std::endl;
            particle_list->read_in_particle_samples_and_filter();
            particle_list->read_in_particle_samples_mixed_event_and_filter();
            const int nev = particle_list->get_number_of_events();
            std::cout << "nev = " << nev << std::endl;
            BF_analysis.calculate_balance_function(particle_list);
            event_id += nev;
        }
        BF_analysis.output_balance_function();
    }
    const double dt =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Program totally finished in " << dt << " sec." << std::endl;
    return 0;
}
