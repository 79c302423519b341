/* Host-only checks of the C++ API (no GPU): configuration reader, run
 * statistics and the output files.  Prints key=value lines for the test. */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>

#include "../../damascus-earth-evap_amd/host/dmsun.hpp"

int main(int argc, char** argv)
{
	const std::string mode = argc > 1 ? argv[1] : "";
	try
	{
		if(mode == "config")
		{
			dmsun::Configuration c(argv[2]);
			std::cout.precision(17);
			std::cout << "ID=" << c.ID << "\nrun_mode=" << c.run_mode << "\ncapture_mode=" << c.capture_mode
			          << "\nsample_size=" << c.sample_size << "\nmax_trajectories=" << c.max_trajectories
			          << "\nfixed_seed=" << c.fixed_seed << "\ninterpolation_points=" << c.interpolation_points
			          << "\nmass_GeV=" << c.DM.mass / dmsun::units::GeV()
			          << "\nsigma_p_cm2=" << c.DM.sigma_proton / (dmsun::units::cm() * dmsun::units::cm())
			          << "\nsigma_e_cm2=" << c.DM.sigma_electron / (dmsun::units::cm() * dmsun::units::cm())
			          << "\ninteraction=" << c.DM.interaction << "\nlow_mass=" << c.DM.low_mass
			          << "\nvobs=" << c.halo.vobs_kms[0] << "," << c.halo.vobs_kms[1] << "," << c.halo.vobs_kms[2]
			          << "\nrho=" << c.halo.rho_GeV_cm3 << "\nv0=" << c.halo.v0_kms << "\nvesc=" << c.halo.vesc_kms
			          << "\nmax_scatterings=" << c.maximum_number_of_scatterings << "\nsnapshot=" << c.snapshot_enabled
			          << "\nwall=" << c.max_trajectory_wall_time_sec << "\n";
		}
		else if(mode == "grid")
		{
			/* tests/test_Solar_Model.cpp:192-224 through the C++ Solar_Model (GPU) */
			dmsun::Solar_Model SSM(0);
			dmsun::DM_Particle DM = dmsun::DM_Particle_SI(0.01 * dmsun::units::GeV());
			DM.Set_Low_Mass_Mode(true);
			DM.Set_Sigma_Proton(1.0e-36 * dmsun::units::cm() * dmsun::units::cm());
			const unsigned int radius_points = 31, speed_points = 29;
			SSM.Interpolate_Total_DM_Scattering_Rate(DM, radius_points, speed_points);
			std::printf("radius_points=%u\nspeed_points=%u\n", SSM.Scattering_Rate_Interpolation_Radius_Points(),
			            SSM.Scattering_Rate_Interpolation_Speed_Points());
			double worst = 0.0;
			for(unsigned int i : {0u, 7u, 15u, 30u})
				for(unsigned int j : {0u, 5u, 14u, 28u})
				{
					const double r = dmsun::units::rSun() * i / (radius_points - 1);
					const double v = 0.75 * j / (speed_points - 1);
					const double want = SSM.Total_DM_Scattering_Rate_Computed(DM, r, v);
					const double got = SSM.Total_DM_Scattering_Rate(DM, r, v);
					/* the reference's tolerance, max(1e-30, 1e-10 |expected|) */
					worst = std::max(worst, std::fabs(got - want) / std::max(1.0e-30, 1.0e-10 * std::fabs(want)));
				}
			std::printf("node_err_over_tolerance=%.17g\n", worst);
			SSM.Interpolate_Total_DM_Scattering_Rate(DM, 1, speed_points);
			const double r = 0.37 * dmsun::units::rSun(), v = 0.21;
			std::printf("off_radius_points=%u\noff_speed_points=%u\noff_equal=%d\n",
			            SSM.Scattering_Rate_Interpolation_Radius_Points(), SSM.Scattering_Rate_Interpolation_Speed_Points(),
			            (int) (SSM.Total_DM_Scattering_Rate(DM, r, v) == SSM.Total_DM_Scattering_Rate_Computed(DM, r, v)));
		}
		else if(mode == "percentile")
		{
			std::vector<double> v;
			for(int i = 3; i < argc; i++)
				v.push_back(std::stod(argv[i]));
			std::printf("p=%.17g\n", dmsun::Percentile(v, std::stod(argv[2])));
		}
		else if(mode == "wilson")
		{
			const auto e = dmsun::Estimate_Binomial_Rate(std::stoull(argv[2]), std::stoull(argv[3]));
			std::printf("rate=%.17g\nse=%.17g\nlo=%.17g\nhi=%.17g\n", e.rate, e.standard_error, e.ci_lower, e.ci_upper);
		}
		else if(mode == "write")
		{
			/* synthetic results: 5 bins, known sums */
			dms_counters c{};
			c.trajectories = 1000;
			c.captured = 150;
			c.completed_outward_escapes = 840;
			c.free_particles = 400;
			c.reflected_particles = 440;
			c.numerical_failures = 6;
			c.computational_truncations = 4;
			c.total_scatterings = 2500;
			c.sync_rounds = 3;
			c.sync_interval = 512;
			c.final_round_trajectories = 100;
			c.early_stop_reason = 1;
			c.computing_time_sec = 2.5;
			const int nb = 5;
			std::vector<double> h(8 * nb);
			for(int q = 0; q < 8; q++)
				for(int b = 0; b < nb; b++)
					h[q * nb + b] = (q % 2 ? 3.0 : 1.0) * (b + 1) * (q < 4 ? 10.0 : 20.0) * ((q / 2) % 2 ? (b + 2.0) : 1.0);
			std::vector<uint64_t> counts(2 * nb, 7);
			std::vector<dms_evap_event> ev(3);
			ev[0] = {0, 0, 11, 1.0, 5.0e7};
			ev[1] = {0, 0, 3, 2.0, 1.0e6};
			ev[2] = {0, 0, 8, 3.0, 5.0e7};
			dmsun::Simulation_Data d(150, 1000);
			d.Set_Results(c, h, counts, ev);
			dmsun::DM_Particle dm = dmsun::DM_Particle_SI(0.3 * dmsun::units::GeV());
			dm.Set_Sigma_Proton(1e-36 * dmsun::units::cm() * dmsun::units::cm());
			d.Write_Output_Files(argv[2], dm);
			d.Print_Summary();
		}
		else
			return 2;
	}
	catch(const std::invalid_argument& e)
	{
		std::cout << "invalid_argument=" << e.what() << "\n";
		return 3;
	}
	catch(const std::exception& e)
	{
		std::cout << "error=" << e.what() << "\n";
		return 4;
	}
	return 0;
}
