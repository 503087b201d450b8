from ott_amd.geometry import costs, epsilon_scheduler, pointcloud

__all__ = ["costs", "epsilon_scheduler", "pointcloud"]
