/*
 * dmsun_run: command-line driver with the reference's run modes
 * "Parameter point" and "Capture" (src/main.cpp:68-99) on one MI355X.
 *
 *   dmsun_run <config.cfg> [--device N] [--bins N] [--solar-model PATH]
 *
 * Reads the configuration (the libconfig keys of the reference), builds the
 * solar model and the rate table on the GPU, runs Generate_Data, prints the
 * reference's summary and, in parameter-point mode, writes bincount.txt and
 * evaporation_times.txt into <output_dir>/results_<log10 m>_<log10 sigma>/.
 */
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "dmsun.hpp"

int main(int argc, char** argv)
{
	if(argc < 2)
	{
		std::cerr << "usage: " << argv[0] << " <config.cfg> [--device N] [--bins N] [--solar-model PATH]" << std::endl;
		return 2;
	}
	int device = 0, bins = 2000;
	std::string model;
	for(int i = 2; i < argc; i++)
	{
		const std::string a = argv[i];
		if(a == "--device" && i + 1 < argc)
			device = std::atoi(argv[++i]);
		else if(a == "--bins" && i + 1 < argc)
			bins = std::atoi(argv[++i]);
		else if(a == "--solar-model" && i + 1 < argc)
			model = argv[++i];
		else
		{
			std::cerr << "unknown option " << a << std::endl;
			return 2;
		}
	}
	try
	{
		dmsun::Configuration cfg(argv[1]);
		cfg.Print_Summary();
		if(cfg.run_mode != "Parameter point" && cfg.run_mode != "Capture")
		{
			std::cerr << "Run mode " << cfg.run_mode << " is outside this build (Parameter point and Capture only)." << std::endl;
			return 1;
		}
		dmsun::Solar_Model SSM(device, model);
		SSM.Print_Summary();
		const double rSun = dmsun::units::rSun();
		dmsun::Simulation_Data data_set(cfg.sample_size, cfg.max_trajectories, 0.0, 1);
		data_set.Configure(2.0 * rSun, 1, cfg.maximum_number_of_scatterings);
		data_set.Set_Bins(bins);
		if(cfg.snapshot_enabled)
			data_set.Set_Snapshots(cfg.output_dir + (cfg.output_dir.empty() || cfg.output_dir.back() == '/' ? "" : "/") + "snapshots",
			                       cfg.snapshot_interval_seconds);
		std::cout << (cfg.capture_mode ? "Generate data in CAPTURE MODE..." : "Generate data...") << std::endl
		          << "\tm_DM [MeV]:\t" << cfg.DM.mass / dmsun::units::MeV() << "\t\tsigma_p [cm2]:\t"
		          << cfg.DM.Sigma_Proton() / (dmsun::units::cm() * dmsun::units::cm()) << std::endl
		          << "\tsigma_e [cm2]:\t" << cfg.DM.Sigma_Electron() / (dmsun::units::cm() * dmsun::units::cm()) << std::endl
		          << std::endl;
		SSM.Interpolate_Total_DM_Scattering_Rate(cfg.DM, cfg.interpolation_points, cfg.interpolation_points);
		data_set.Generate_Data(cfg.DM, SSM, cfg.halo, cfg.fixed_seed, cfg.capture_mode);
		if(cfg.capture_mode)
			data_set.Print_Capture_Mode_Summary();
		else
			data_set.Print_Summary();
		if(!cfg.capture_mode)
		{
			const std::string prefix = "results_";
			std::string dir = cfg.output_dir;
			if(!dir.empty() && dir.back() != '/')
				dir += "/";
			dir += prefix + std::to_string(std::log10(cfg.DM.mass / dmsun::units::GeV())) + "_"
			       + std::to_string(std::log10(cfg.DM.Sigma_Proton() / (dmsun::units::cm() * dmsun::units::cm()))) + "/";
			data_set.Write_Output_Files(dir, cfg.DM);
			std::cout << "Output written to " << dir << std::endl;
		}
	}
	catch(const std::exception& e)
	{
		std::cerr << "Error: " << e.what() << std::endl;
		return 1;
	}
	return 0;
}
