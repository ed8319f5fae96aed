"""saddlepoint_amd: an MI355X-native backend for the saddlepoint PDHG LP solver."""
